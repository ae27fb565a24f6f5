/*
 * vedas_hip.h -- C ABI of the MI355X (gfx950) VEDAS query-path library
 * (libvedas_hip.so).
 *
 * Plain C: opaque handles, raw device pointers and sizes, int status codes.
 * Every entry point returns VD_OK (0) or a negative VD_E* code; the message
 * of the last failure on a context is returned by vd_ctx_last_error().
 * All device work of a context is issued on that context's single HIP
 * stream, in call order.  Calls that return a size to the host (join,
 * select, unique) synchronise the stream once to read it back.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference tree Remixman/Vedas).  The reference
 * runs these steps through Thrust and ModernGPU on CUDA; here they are
 * hand-written HIP kernels.
 */
#ifndef VEDAS_HIP_H
#define VEDAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define VD_OK       0
#define VD_EINVAL  (-1)   /* bad argument, shape or handle                 */
#define VD_ENOMEM  (-2)   /* device or host allocation failed              */
#define VD_EHIP    (-3)   /* a HIP runtime call failed                     */
#define VD_ERANGE  (-4)   /* size beyond what the kernels support          */

#define VD_MAX_COLS 16    /* columns per relation (the reference caps at 9) */
#define VD_UNBOUND  0xFFFFFFFFu  /* "no constant" marker in batched lookups */

typedef struct vd_ctx   vd_ctx;    /* device + stream + allocator + profiler */
typedef struct vd_rel   vd_rel;    /* column-store relation of uint32 IDs    */
typedef struct vd_index vd_index;  /* HBM-resident triple permutations       */
typedef struct vd_graph vd_graph;  /* a captured run, replayable as one launch */

/* Permutations.  XYZ = sort order; X is implicit through the L1 index, Y and
 * Z are stored as columns (the reference's xy_data / xyz_data arrays,
 * VedasStorage.h:87-95). */
enum {
    VD_PERM_SPO = 0, VD_PERM_SOP = 1, VD_PERM_PSO = 2,
    VD_PERM_POS = 3, VD_PERM_OSP = 4, VD_PERM_OPS = 5,
    VD_NUM_PERMS = 6
};
#define VD_PERM_BIT(p) (1u << (p))

/* Device view of one permutation.  Sizes are element counts. */
typedef struct {
    uint64_t n;               /* triples                                   */
    uint64_t n_l1;            /* distinct X                                */
    uint64_t n_l2;            /* distinct (X,Y)                            */
    const uint32_t *l1_values, *l1_offsets;   /* distinct X, first row     */
    const uint32_t *l2_values, *l2_offsets;   /* Y of each (X,Y), first row */
    const uint32_t *col_y, *col_z;            /* Y and Z per row, sorted   */
} vd_perm_view_t;

/* Inclusive value interval [lo, hi] removed from a select (the reference's
 * empty-interval filter, EmptyIntervalDict.cu:108-170). */
typedef struct { uint32_t lo, hi; } vd_interval_t;

/* ---- context ----------------------------------------------------------- */
/* Replaces mgpu::standard_context_t (VedasQuery.cu:86) plus the implicit
 * Thrust device state: one per host thread / request slot. */
int  vd_ctx_create(int device, vd_ctx **out);
int  vd_ctx_destroy(vd_ctx *ctx);
const char *vd_ctx_last_error(const vd_ctx *ctx);
int  vd_ctx_sync(vd_ctx *ctx);
void *vd_ctx_stream(vd_ctx *ctx);             /* the hipStream_t          */
int  vd_ctx_trim(vd_ctx *ctx);                /* release cached blocks    */
/* Device memory in use / cached by the context's allocator, in bytes. */
int  vd_ctx_mem(vd_ctx *ctx, uint64_t *in_use, uint64_t *cached);
/* Highest in-use byte count since the last reset (reset != 0 restarts it
 * from the current in-use count): the measured peak that admission
 * estimates are compared with. */
int  vd_ctx_mem_peak(vd_ctx *ctx, uint64_t *peak, int reset);
/* Tunables: "radix_items" (keys per thread in a radix tile: 16 = 512
 * threads x 16 keys, the default; 8 = 512 x 8; 24 = 256 x 24, keys only),
 * "radix_bits" (digit width of the IndexSwap re-sort: 0 = 9-bit digits when
 * they save a pass, else 8-bit, the default; 8 = always 8-bit; 9 = 9-bit
 * whenever the key fits 27 bits), "radix_group" (radix tiles an XCD takes
 * as one group of consecutive tiles: 8, the default, or 4, 16, 32; 0 = one
 * global tile order), "drive_min_rows" (smallest input rows, both inputs
 * together, of a driven join: 2^20 by default), "join_nv" (merge-path
 * diagonal per join
 * expansion block), "join_rcap" (right keys staged in LDS per bounds
 * block).  Returns VD_EINVAL for an unknown key or an unsupported value. */
int  vd_ctx_tune(vd_ctx *ctx, const char *key, int value);

/* ---- graph capture (no counterpart in the reference) ---------------------
 * A run of calls on a context can be captured into one HIP graph and
 * relaunched without host work between its kernels.  Sizes the host reads
 * back decide allocations and launch shapes, so a run is captured in two
 * steps: run it once with vd_ctx_record(ctx, 1) (every read-back word is
 * logged), then again between vd_ctx_capture_begin and vd_ctx_capture_end
 * (nothing executes: read-backs return the logged words, in order, and the
 * kernels and copies go into the graph).  The second run must make the same
 * calls on the same data.  Device memory the capture allocates or frees
 * belongs to the graph until vd_graph_free (relations made during the
 * capture read and hold it; destroying them frees nothing); host data
 * uploaded during the capture is copied into buffers of the graph.
 * vd_copy_h2d / vd_copy_d2h are refused during a capture.  Every launch of
 * the graph redoes the captured work and rewrites those relations' columns.
 * vd_ctx_capture_end with out == NULL discards the capture (use it when the
 * captured run fails).  Free every graph of a context before the context. */
int  vd_ctx_record(vd_ctx *ctx, int on);
int  vd_ctx_capture_begin(vd_ctx *ctx);
int  vd_ctx_capture_end(vd_ctx *ctx, vd_graph **out);
int  vd_graph_launch(vd_ctx *ctx, vd_graph *g);
int  vd_graph_free(vd_ctx *ctx, vd_graph *g);

/* ---- device buffers ---------------------------------------------------- */
int  vd_dev_alloc(vd_ctx *ctx, size_t bytes, void **dptr);
int  vd_dev_free(vd_ctx *ctx, void *dptr);
int  vd_copy_h2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes);
int  vd_copy_d2h(vd_ctx *ctx, void *dst, const void *src, size_t bytes);
int  vd_copy_d2d(vd_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- index build (VedasStorage) ---------------------------------------- */
/* Replaces VedasStorage::VedasStorage(RdfData&, preload, fullIndex)
 * (VedasStorage.cu:25-73) with createIndex1Level / createIndex
 * (VedasStorage.cu:270-358): for each permutation in perm_mask, sort the
 * (X,Y,Z) triples lexicographically (duplicates kept) and build
 * L1 = distinct X + first row, L2 = distinct (X,Y) + Y + first row.
 * d_s/d_p/d_o are device arrays of n uint32 IDs; they are not modified.
 * The index stays resident in HBM until vd_index_free. */
int  vd_index_build(vd_ctx *ctx, const uint32_t *d_s, const uint32_t *d_p,
                    const uint32_t *d_o, uint64_t n, uint32_t perm_mask,
                    vd_index **out);
/* An index over permutation arrays already in device memory (what a .vds
 * file holds: VedasStorage::open, VedasStorage.cu:452-484): for every
 * permutation p in perm_mask, views[p] gives n, n_l1, n_l2 and the six
 * arrays with vd_index_view's meaning.  The arrays are copied (the caller
 * keeps its buffers); only sizes are checked, not the sort order.  Two
 * arrays the reference's full-index file does not hold may be left NULL
 * and are then rebuilt on the device: col_y (each row's Y from its L2 run;
 * the file has no Y column for SPO, SOP and OPS) or l2_values / l2_offsets
 * (the (X, Y) run heads of the rows; the file has no L2 for OSP). */
int  vd_index_from_arrays(vd_ctx *ctx, uint32_t perm_mask, const vd_perm_view_t *views,
                          vd_index **out);
/* Rows of runs: out[r] = values[k] for the run k that holds row r, run k
 * covering rows [offsets[k], offsets[k+1]) and the last run ending at n
 * (device arrays).  Expands an L1 or L2 array into its column, e.g. the
 * predicate of every PSO row to rebuild the triples of a .vds file. */
int  vd_runs_expand(vd_ctx *ctx, const uint32_t *values, const uint32_t *offsets,
                    uint64_t nruns, uint64_t n, uint32_t *out);
int  vd_index_free(vd_ctx *ctx, vd_index *idx);
int  vd_index_view(const vd_index *idx, int perm, vd_perm_view_t *out);
uint32_t vd_index_perms(const vd_index *idx);   /* mask of built perms */

/* Row range [begin, end) of the rows whose X == k1 (and Y == k2 when
 * has_k2).  Replaces QueryExecutor::findL2OffsetFromL1 and
 * findDataOffsetFromL2 (QueryExecutor.cu:596-646), with an absent constant
 * giving an empty range (the reference returns the next key's range). */
int  vd_index_lookup(vd_ctx *ctx, const vd_index *idx, int perm, uint32_t k1,
                     int has_k2, uint32_t k2, uint64_t *begin, uint64_t *end);
/* Statistics of a permutation for size estimates: for every L1 entry i
 * (leading key X = l1_values[i]) the number of rows and the number of
 * distinct (X,Y) pairs, into device arrays of n_l1 entries.  Plays the role
 * of the reference's per-key run lengths and Histogram / DataMetaInfo
 * statistics (VedasStorage.cu:278-288,342-355; Histogram.cu:170-231). */
int  vd_index_stats(vd_ctx *ctx, const vd_index *idx, int perm,
                    uint64_t *d_rows, uint64_t *d_pairs);
/* Equal-width and equal-depth histograms of a permutation's leading column
 * (subjects for SPO / SOP, objects for OSP / OPS), as
 * Histogram::initEqWidth / initEqDepth build them from the sorted column at
 * index build time (Histogram.cu:170-231, VedasStorage.cu:316-317,
 * 518-536).  Only rows whose key is below id_limit count (the reference
 * passes LITERAL_START_ID - 1: literal objects left out; 0xFFFFFFFF keeps
 * every key).  VD_HIST_EQUAL_WIDTH: param is the width in IDs; bucket b
 * holds keys [first + b*param, first + (b+1)*param - 1] up to the last key
 * (the reference drops the last key when the span is a multiple of the
 * width).  VD_HIST_EQUAL_DEPTH: param is the depth in rows; a bucket starts
 * at a row, ends at the row param rows later (or the last row) extended to
 * the end of that key's run, so it holds param + 1 rows or more; lower /
 * upper are its first and last key.  lowers / uppers / freqs are host
 * arrays of cap entries; *nbuckets receives the bucket count (VD_ERANGE,
 * nothing written, when it exceeds cap).  Computed from the L1 arrays on
 * the device; no row is read. */
#define VD_HIST_EQUAL_WIDTH 0
#define VD_HIST_EQUAL_DEPTH 1
int  vd_index_histogram(vd_ctx *ctx, const vd_index *idx, int perm, int kind,
                        uint32_t param, uint32_t id_limit, uint32_t *lowers,
                        uint32_t *uppers, uint64_t *freqs, uint64_t cap,
                        uint64_t *nbuckets);
/* Batched form on device arrays: d_k2 may be NULL (one constant for every
 * query) or hold VD_UNBOUND for queries with one constant. */
int  vd_index_lookup_batch(vd_ctx *ctx, const vd_index *idx, int perm,
                           const uint32_t *d_k1, const uint32_t *d_k2,
                           uint64_t nq, uint64_t *d_begin, uint64_t *d_end);

/* ---- relations (FullRelationIR) ---------------------------------------- */
/* Replaces FullRelationIR(columnNum, relationSize) and its column accessors
 * (FullRelationIR.cu:19-41, 524-592, 715-745).  Columns are separate device
 * arrays of nrows uint32; column 0 is the sort key.  Relations produced by
 * vd_select / vd_select_multi / vd_join / vd_join_multi carry an upper bound
 * of each column's values (from the index's column maxima), which
 * vd_sort_col0 uses as its digit plan instead of measuring the maximum;
 * vd_rel_col hands the column out writable and drops that bound. */
int  vd_rel_create(vd_ctx *ctx, int ncols, uint64_t nrows, vd_rel **out);
int  vd_rel_destroy(vd_ctx *ctx, vd_rel *rel);
int  vd_rel_ncols(const vd_rel *rel);
uint64_t vd_rel_nrows(const vd_rel *rel);
uint32_t *vd_rel_col(vd_rel *rel, int c);
int  vd_rel_swap_cols(vd_rel *rel, int i, int j);   /* swapColumn :715  */
int  vd_rel_drop_col(vd_ctx *ctx, vd_rel *rel, int i); /* removeColumn :729 */
/* No counterpart in the reference, whose selects always copy.  A column of
 * a relation returned by an unfiltered vd_select / vd_select_multi is a
 * read-only view of the index permutation (no copy); it keeps the index
 * arrays alive after vd_index_free until the relation is destroyed.  The
 * operators never write into a view (they replace the column); a caller
 * that wants to write through vd_rel_col first calls vd_rel_materialize,
 * which gives every view column a private copy.  vd_rel_is_view(rel, c) is
 * 1 for a view column, else 0. */
int  vd_rel_materialize(vd_ctx *ctx, vd_rel *rel);
int  vd_rel_is_view(const vd_rel *rel, int c);
/* First and last value of column 0 as vd_join computed them with its size
 * read-back: returns 1 and sets *lo and *hi, or 0 when unknown (another
 * producer, column 0 changed since, or no rows).  Replaces the two device
 * reads of the join result's ends in JoinQueryJob::join
 * (JoinQueryJob.cu:388-397). */
int  vd_rel_key_range(const vd_rel *rel, uint32_t *lo, uint32_t *hi);

/* ---- select (SelectQueryJob) -------------------------------------------- */
/* Rows [begin, end) of a permutation (a view when nothing filters them,
 * a compacted copy otherwise; see vd_rel_materialize).  cols: 1 = Z only (the
 * 1-variable select, SelectQueryJob.cu:92-121), 2 = Y only,
 * 3 = Y then Z (2-variable select, :122-264).  With clip, rows whose Y is
 * outside [lo, hi] are dropped (variable bound, :130-141); rows whose Y lies
 * in one of the nskip intervals are dropped (:170-246); with distinct and
 * cols == 2 adjacent duplicates of Y are removed (:151-158).  clip, skip and
 * distinct need Y as column 0 (cols 2 or 3) and a range inside one X group. */
int  vd_select(vd_ctx *ctx, const vd_index *idx, int perm,
               uint64_t begin, uint64_t end, int cols,
               int clip, uint32_t lo, uint32_t hi,
               const vd_interval_t *skip, int nskip, int distinct,
               vd_rel **out);
/* vd_index_lookup + vd_select in one device step with one read-back: the
 * rows of key (k1[, k2]) in the permutation, then filtered as vd_select.
 * Replaces the findL2OffsetFromL1 / findDataOffsetFromL2 lookup followed
 * by SelectQueryJob::startJob (QueryExecutor.cu:596-646,
 * SelectQueryJob.cu:86-270).  begin/end (may be NULL) receive the range. */
int  vd_select_key(vd_ctx *ctx, const vd_index *idx, int perm, uint32_t k1,
                   int has_k2, uint32_t k2, int cols, int clip, uint32_t lo,
                   uint32_t hi, const vd_interval_t *skip, int nskip,
                   int distinct, vd_rel **out, uint64_t *begin, uint64_t *end);
/* Batched range select: lookups for nq queries, then every matching row of
 * every query written back to back as (Y, Z) pairs into d_out_y / d_out_z
 * with d_row_offsets[q] = first output row of query q (nq + 1 entries).
 * When the rows exceed max_rows nothing is written, *total_rows holds the
 * size needed and VD_ERANGE is returned.  max_rows == 0 only counts. */
int  vd_select_batch(vd_ctx *ctx, const vd_index *idx, int perm,
                     const uint32_t *d_k1, const uint32_t *d_k2, uint64_t nq,
                     uint64_t *d_row_offsets, uint32_t *d_out_y,
                     uint32_t *d_out_z, uint64_t max_rows,
                     uint64_t *total_rows);

/* No counterpart in the reference.  vd_select_batch for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it): the same arguments and results, but
 * the per-query first rows, row counts and the row-offset scan's partial
 * sums live in one fixed device array of the library (16 MiB per GPU, held
 * until the process ends) instead of allocations, so at most
 * VD_NOALLOC_MAX_QUERIES queries per call (VD_ERANGE beyond).  It calls no
 * hipMalloc / hipFree.  Calls that overlap on one GPU share that array and
 * race; serialise them (one stream, or one call at a time). */
#define VD_NOALLOC_MAX_QUERIES (1u << 20)
int  vd_select_batch_noalloc(vd_ctx *ctx, const vd_index *idx, int perm,
                             const uint32_t *d_k1, const uint32_t *d_k2,
                             uint64_t nq, uint64_t *d_row_offsets,
                             uint32_t *d_out_y, uint32_t *d_out_z,
                             uint64_t max_rows, uint64_t *total_rows);

/* ---- join (JoinQueryJob) ------------------------------------------------ */
/* Bag inner sort-merge join of L and R on their first nkeys (1 or 2)
 * columns, both sorted ascending by those columns.  For each left row in
 * order, the matching right rows in order (left-major, as in
 * mgpu::inner_join, JoinQueryJob.cu:68-101, 262-269).  Output columns:
 * [key columns, L columns after the keys, R columns after the keys]
 * (JoinQueryJob.cu:287-292, 322-380).  Output column 0 stays sorted.
 * Input relations are not modified. */
int  vd_join(vd_ctx *ctx, const vd_rel *L, const vd_rel *R, int nkeys,
             vd_rel **out);
/* 1 when the rows are known to be pairwise distinct, else 0.  Selects of
 * an index built without repeated triples return distinct rows (the build
 * counts repeated rows of its sorted keys); joins of distinct inputs and
 * the operators that only reorder or drop rows keep the property; dropping
 * a column or vd_rel_col clears it.  vd_unique_adjacent (and _multi) leave
 * such a relation as it is without reading it. */
int  vd_rel_distinct(const vd_rel *rel);

/* Join driven by a relation in output order (no single counterpart in the
 * reference: it yields what JoinQueryJob::join followed by
 * IndexSwapJob::startJob, JoinQueryJob.cu:262-380 + IndexSwapJob.cu:28-95,
 * yield when D is the re-indexed copy of one input, e.g. the POS range of a
 * pattern whose PSO range was selected).  For every row of D, in D's
 * order, the rows of Y whose column 0 equals D's column dkey, in Y's order.
 * Y is sorted ascending on column 0; D may be in any order.  D rows whose
 * key is outside [key_lo, key_hi] match nothing.  Output column c is Y's
 * column src[c] (src[c] >= 0) or D's column -1 - src[c].  *match_lo /
 * *match_hi receive the smallest and largest key that matched.
 * Returns VD_OK with *out == NULL when the driven form does not apply: D
 * and Y hold fewer rows together than the "drive_min_rows" tunable (2^20
 * by default), or the key values (bounded by
 * key_hi and the columns' value bounds) exceed 2^28 - 1 or twice the rows
 * of D and Y (the run table of Y is zero-filled over every key value, 8
 * bytes each), or, with check_dups, two adjacent equal D rows match more
 * than one row each (the join-then-sort order would interleave them
 * differently). */
int  vd_join_driven(vd_ctx *ctx, const vd_rel *D, int dkey, const vd_rel *Y,
                    uint32_t key_lo, uint32_t key_hi, int nout, const int *src,
                    int check_dups, vd_rel **out, uint32_t *match_lo,
                    uint32_t *match_hi);
/* 1 when vd_join_driven(D, dkey, Y, .., key_hi, ..) takes the driven form
 * (its size rules, which do not depend on row order), else 0: a caller
 * about to re-sort D for it asks first. */
int  vd_join_driven_applies(const vd_ctx *ctx, const vd_rel *D, int dkey,
                            const vd_rel *Y, uint32_t key_hi);
/* vd_join(L, R, nkeys), unless its output would have more than max_rows
 * rows: then *out = NULL and nothing is built (the count phase of
 * JoinQueryJob::join, JoinQueryJob.cu:262-269, has run).  *rows receives
 * the output size either way.  For a caller that takes another form for
 * large outputs (the driven join after re-sorting its left input). */
int  vd_join_capped(vd_ctx *ctx, const vd_rel *L, const vd_rel *R, int nkeys,
                    uint64_t max_rows, vd_rel **out, uint64_t *rows);

/* ---- re-sort and finalise (FullRelationIR / IndexSwapJob / QueryPlan) --- */
/* Stable sort of all rows by column 0 only.  Replaces
 * FullRelationIR::sortByFirstColumn (FullRelationIR.cu:594-713) as used by
 * IndexSwapJob::startJob (IndexSwapJob.cu:28-95). */
int  vd_sort_col0(vd_ctx *ctx, vd_rel *rel);
/* The rows of X whose column xcol value occurs in column ycol of Y, in X's
 * order (columns, bounds and the distinct mark of X).  For an IndexSwap
 * (IndexSwapJob.cu:28-95) whose output only feeds a join on the swapped
 * column (JoinQueryJob.cu:125-453): the join of the stably sorted X keeps
 * exactly these rows, in the same order, so filtering first and sorting
 * the rows kept gives the same join output.  Keys above key_hi (or a known
 * bound of either column) match nothing; *out = NULL with *rows = ~0 when
 * that bound exceeds 2^28 - 1 (the key bitmap's size), and *out = NULL with
 * *rows = the kept count when it exceeds max_rows (nothing is written). */
int  vd_semi_filter(vd_ctx *ctx, const vd_rel *X, int xcol, const vd_rel *Y, int ycol,
                    uint32_t key_hi, uint64_t max_rows, vd_rel **out, uint64_t *rows);
/* Lexicographic sort of all rows over all columns.  Replaces
 * FullRelationIR::sort (FullRelationIR.cu:218-336). */
int  vd_sort_lex(vd_ctx *ctx, vd_rel *rel);
/* Drop rows equal (in every column) to the row before them; no sort.
 * Replaces FullRelationIR::removeDuplicate (FullRelationIR.cu:338-522). */
int  vd_unique_adjacent(vd_ctx *ctx, vd_rel *rel);

/* ---- batched operators (many queries per launch) ------------------------ */
/* No counterpart in the reference, which runs one query and one operator at
 * a time (QueryPlan.cu:160-205).  Each call runs one operator for many
 * independent inputs with a fixed number of kernel launches and one stream
 * synchronisation, whatever the number of inputs; results equal those of
 * the single-input entry points above, input by input. */
typedef struct {
    int      perm;   /* VD_PERM_*                                          */
    uint32_t k1;     /* X constant                                          */
    uint32_t k2;     /* Y constant, or VD_UNBOUND                           */
    int      cols;   /* 1 = Z only, 3 = (Y, Z)   (as vd_select)            */
} vd_select_job_t;
/* outs[i] = vd_select of job i over its lookup range (no clip/skip): a
 * view of the permutation's rows. */
int  vd_select_multi(vd_ctx *ctx, const vd_index *idx,
                     const vd_select_job_t *jobs, int njobs, vd_rel **outs);
/* begin[i]/end[i] = vd_index_lookup of job i (its perm, k1 and k2; cols is
 * not used), for every job in one kernel launch and one read-back.  The
 * plan executor resolves all of a plan's select ranges with it before the
 * first operator runs; in the reference each SelectQueryJob does its own
 * findL2OffsetFromL1 / findDataOffsetFromL2 (QueryExecutor.cu:596-646). */
int  vd_index_lookup_multi(vd_ctx *ctx, const vd_index *idx,
                           const vd_select_job_t *jobs, int njobs,
                           uint64_t *begin, uint64_t *end);
/* vd_index_lookup_multi plus, for every job, the first and last value of
 * the column after its constants over the rows found: the Y column for a
 * job with k1 only, the Z column for a job with k1 and k2 (cols is not
 * used).  That column is sorted inside the range, so first/last are its
 * minimum and maximum (both 0 for an empty range).  Replaces the host
 * binary searches and data reads of the bound pre-scan,
 * QueryExecutor::updateDataBound / updateL2Bound (QueryExecutor.cu:
 * 648-675, 737-765), for all patterns of a query in one launch. */
int  vd_index_bounds_multi(vd_ctx *ctx, const vd_index *idx,
                           const vd_select_job_t *jobs, int njobs,
                           uint64_t *begin, uint64_t *end,
                           uint32_t *first, uint32_t *last);
/* outs[i] = vd_join(L[i], R[i], 1). */
int  vd_join_multi(vd_ctx *ctx, const vd_rel *const *L, const vd_rel *const *R,
                   int npairs, vd_rel **outs);
/* vd_sort_col0 on every relation. */
int  vd_sort_col0_multi(vd_ctx *ctx, vd_rel *const *rels, int n);
/* vd_unique_adjacent on every relation. */
int  vd_unique_adjacent_multi(vd_ctx *ctx, vd_rel *const *rels, int n);

/* ---- synthetic triples (host) ------------------------------------------ */
/* Seeded generator of (S,P,O) IDs: subjects over [1, n_entities] uniform
 * (subject_zipf == 0) or Zipf(subject_zipf) by rank; predicates over
 * [1, n_predicates] Zipf(pred_zipf) with ID = rank; objects an entity
 * (uniform over [1, n_entities]) with probability 1 - literal_frac, else a
 * literal literal_base + U[0, n_literals).  With distinct, repeated triples
 * are redrawn until n distinct triples exist.  Host arrays of n entries. */
typedef struct {
    uint64_t seed;
    uint64_t n;
    uint32_t n_entities;
    double   subject_zipf;     /* 0 = uniform                               */
    uint32_t n_predicates;
    double   pred_zipf;
    double   literal_frac;
    uint32_t literal_base;
    uint32_t n_literals;
    int      distinct;
} vd_gen_params_t;
int  vd_gen_triples(const vd_gen_params_t *prm, uint32_t *s, uint32_t *p,
                    uint32_t *o);

/* ---- profiling --------------------------------------------------------- */
/* When enabled, every kernel launch is bracketed by HIP events on the
 * context stream; per-kernel launch counts, device time and algorithmic
 * HBM bytes are folded in at the next vd_ctx_sync. */
int  vd_prof_enable(vd_ctx *ctx, int on);
/* Restricts the timing to the launches whose label is in the
 * comma-separated list (NULL or "" = every launch): each timed launch adds
 * two events to the stream, so a caller timing one kernel class keeps the
 * others free of that cost. */
int  vd_prof_only(vd_ctx *ctx, const char *labels);
int  vd_prof_reset(vd_ctx *ctx);
int  vd_prof_count(vd_ctx *ctx);
int  vd_prof_get(vd_ctx *ctx, int i, const char **name, uint64_t *launches,
                 double *ms, double *bytes);

#ifdef __cplusplus
}
#endif
#endif /* VEDAS_HIP_H */
