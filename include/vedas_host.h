/*
 * vedas_host.h -- C entry points of libvedas.so, the C++ operator API
 * (VedasStorage, SelectQueryJob, JoinQueryJob, IndexSwapJob, QueryPlan,
 * QueryExecutor).  They run one query through that API the way the
 * reference's vdQuery does (VedasQuery.cu:136-205 -> QueryExecutor::query),
 * for callers that cannot link C++ (Python via ctypes, other FFIs).
 *
 * Query text:  SELECT <* | ?v ...> WHERE { s p o . s p o ... } PLAN <node>
 *   terms: ?name for a variable, a decimal ID for a constant;
 *   node:  (U <pattern index> <var>)        select (UPLOAD), var = the
 *                                           variable the index is chosen for
 *          (J <var> <node> <node>)          join on var
 *          (S <var> <node>)                 index swap to var
 *   Without PLAN, the query graph planner (QueryGraph::generateQueryPlan,
 *   QueryGraph.cu:64-279) builds the plan from each pattern's index-range
 *   rows (the reference's estimate_rows, QueryExecutor.cu:666,758).
 */
#ifndef VEDAS_HOST_H
#define VEDAS_HOST_H

#include <stdint.h>

#include "vedas_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vdq_engine vdq_engine;   /* context + resident storage */
typedef struct vdq_result vdq_result;   /* final relation + statistics */

typedef struct {
    uint64_t rows;          /* final rows                               */
    uint64_t join_rows;     /* sum of join output rows                  */
    uint64_t select_rows;   /* sum of select output rows                */
    uint64_t swap_rows;     /* rows re-sorted by index swaps            */
    double elapsed_ms;      /* host wall time of the query              */
    double join_ms, select_ms, swap_ms;
    int jobs;               /* operators executed                       */
    int driven_joins;       /* joins that delivered their IndexSwap's order
                               (those swaps re-sort nothing)             */
    uint64_t filtered_rows; /* rows an IndexSwap left to its join dropped
                               before the sort (they match nothing)      */
} vdq_stats_t;

int vdq_engine_create(int device, vdq_engine **out);
int vdq_engine_free(vdq_engine *e);
/* Build the store from host arrays (perm_mask of VD_PERM_BIT). */
int vdq_engine_load(vdq_engine *e, const uint32_t *s, const uint32_t *p,
                    const uint32_t *o, uint64_t n, uint32_t perm_mask);
/* Build the store from device arrays of the engine's context. */
int vdq_engine_load_device(vdq_engine *e, const uint32_t *d_s,
                           const uint32_t *d_p, const uint32_t *d_o,
                           uint64_t n, uint32_t perm_mask);
/* Front-ends (no counterpart API in the reference, which reads N-Triples
 * with raptor2 and SPARQL with rasqal: loader.cu:12-93, SparqlQuery.cu:
 * 29-124).  A dictionary numbers the terms of an N-Triples file as the
 * reference's loader does (subjects/objects from 1, predicates from 1,
 * literal objects from 5e7 with literal_dict); vdq_engine_load_dict builds
 * the engine's store from its triples and keeps a pointer to the
 * dictionary (the caller keeps it alive).  vdq_sparql runs a SPARQL
 * SELECT over a basic graph pattern (PREFIX, SELECT vars|*, WHERE {
 * triples with ; and , }) by mapping its terms to IDs (unknown terms match
 * nothing) and planning it with the query graph planner; vdq_dict_term
 * turns result IDs back into terms (predicate columns:
 * vdq_result_is_predicate). */
typedef struct vdq_dict vdq_dict;
int vdq_dict_create(int literal_dict, vdq_dict **out);
int vdq_dict_free(vdq_dict *d);
int vdq_dict_load_ntriples(vdq_dict *d, const char *path, uint64_t *n_triples);
int vdq_dict_load_ntriples_text(vdq_dict *d, const char *text, uint64_t *n_triples);
int vdq_dict_triples(const vdq_dict *d, uint32_t *s, uint32_t *p, uint32_t *o);
int vdq_dict_term(const vdq_dict *d, int is_predicate, uint32_t id, char *out, size_t cap);
int vdq_dict_translate(const vdq_dict *d, const char *sparql, char *out, size_t cap);
int vdq_engine_load_dict(vdq_engine *e, const vdq_dict *d, uint32_t perm_mask);
/* The .vdd dictionary file of a store (save_dict, VedasBuild.cu:145-160;
 * load_dict, VedasQuery.cu:213-268): per dictionary (subject/object,
 * predicate, literal) a count line, then "id term" lines, terms in
 * N-Triples form.  vdq_dict_open replaces the dictionary's terms (its
 * triples are cleared).  vdq_engine_set_dict gives an engine whose store
 * came from a .vds file (vdq_engine_open) the dictionary vdq_sparql uses
 * (the caller keeps it alive). */
int vdq_dict_save(const vdq_dict *d, const char *path);
int vdq_dict_open(vdq_dict *d, const char *path);
int vdq_engine_set_dict(vdq_engine *e, const vdq_dict *d);
int vdq_sparql(vdq_engine *e, const char *sparql, vdq_result **out);
int vdq_result_is_predicate(const vdq_result *r, int col);

/* The .vds store file: VedasStorage::write / VedasStorage::open
 * (VedasStorage.cu:452-516).  save writes the default index (PSO, POS) and,
 * when SPO, SOP, OSP and OPS are built too, the full-index arrays; open
 * replaces the engine's store with the file's permutations (no sort; a
 * full-index file written by the reference, whose subject L1 is empty, gets
 * SPO and SOP by an index build over its triples). */
int vdq_engine_save(vdq_engine *e, const char *path);
int vdq_engine_open(vdq_engine *e, const char *path);
vd_ctx *vdq_engine_ctx(vdq_engine *e);
vd_index *vdq_engine_index(vdq_engine *e);

int vdq_query(vdq_engine *e, const char *text, vdq_result **out);

/* The plan of a query text as a PLAN clause: the text's own, or the query
 * graph planner's with est_rows[i] as pattern i's range rows (no device
 * work).  Replaces QueryGraph(...).generateQueryPlan() (QueryGraph.cu:
 * 38-74).  VD_ERANGE if the plan does not fit cap bytes. */
int vdq_plan(const char *text, const uint64_t *est_rows, char *out, size_t cap);

/* Autotuner (no counterpart in the reference, whose kernel shapes are
 * fixed by ModernGPU and Thrust): times the given queries under each
 * candidate of the kernel tunables -- radix tile shape ("radix_items"),
 * merge-path diagonal ("join_nv"), LDS right-key capacity ("join_rcap"),
 * radix digit width of the re-sorts ("radix_bits"), radix tiles per XCD
 * group ("radix_group"), see vd_ctx_tune -- by
 * coordinate descent, keeps the fastest and applies it to the engine's
 * contexts (serving slots included).  Results do not depend on the
 * tunables. */
typedef struct {
    int radix_items, join_nv, join_rcap;  /* chosen values               */
    int radix_bits;                       /* IndexSwap digit width (0 = by rule) */
    double base_ms;                       /* queries with the defaults    */
    double tuned_ms;                      /* queries with the chosen set  */
    int evaluations;
    int radix_group;                      /* radix tiles per XCD group (0 = one order) */
} vdq_tune_t;
int vdq_autotune(vdq_engine *e, const char *const *texts, int n, int reps, vdq_tune_t *out);

/* Many queries at once through the batched operators (no counterpart in
 * the reference): all plans advance together in rounds, one batched call
 * per operator kind per round.  outs[i] is query i's result, equal to
 * vdq_query's after ordering rows and columns canonically. */
typedef struct {
    int rounds;             /* operator rounds                          */
    int calls;              /* batched C-ABI operator calls             */
    uint64_t join_rows, select_rows, swap_rows;
    double elapsed_ms;      /* host wall time of the whole batch        */
} vdq_batch_stats_t;
int vdq_batch(vdq_engine *e, const char *const *texts, int n, vdq_result **outs,
              vdq_batch_stats_t *stats);

/* A batch captured as one HIP graph (no counterpart in the reference,
 * whose QueryPlan::execute issues every operator from the host,
 * QueryPlan.cu:160-205): vdq_batch's work is run once with every size the
 * host reads back logged, run again under stream capture with those sizes
 * (vd_ctx_record / vd_ctx_capture_*), and the graph is launched once to
 * fill the results.  vdq_graph_replay relaunches it and waits: the same
 * kernels over the same store with no host work between them, the results
 * rewritten in place.  For a batch that is run again and again (a fixed
 * set of queries over a store that does not change); a replay after the
 * store changed is wrong.  vdq_graph_result gives result i, owned by the
 * graph (do not free it; valid until vdq_graph_free).  The graph holds the
 * batch's device memory; free it before its engine. */
typedef struct vdq_graph vdq_graph;
int vdq_graph_create(vdq_engine *e, const char *const *texts, int n, vdq_graph **out);
int vdq_graph_replay(vdq_graph *g);
vdq_result *vdq_graph_result(vdq_graph *g, int i);
int vdq_graph_free(vdq_graph *g);
/* Serving (no counterpart in the reference): requests arrive open loop at
 * arrival_ms (relative to the call), are admitted while their estimated
 * peak memory fits mem_budget, and run in continuous batches on `slots`
 * concurrent streams.  plain != 0 runs each request alone through the
 * single-query path on one stream (the baseline).  recs[i] reports
 * request i; outs (may be NULL) receives the results.  Fill policy: while
 * two or more batches are running, a free slot starts the next one only
 * once min_fill requests are queued or the oldest has waited max_wait_ms
 * (min_fill 0: start at once); otherwise it starts at once. */
typedef struct {
    int slots;
    int max_batch;
    uint64_t mem_budget;
    int plain;
    int min_fill;
    double max_wait_ms;
} vdq_serve_cfg_t;
typedef struct {
    double arrival_ms, start_ms, done_ms;
    double est_bytes;        /* admission estimate of its peak bytes      */
    uint64_t rows;
    int batch, slot;
    int mem_held;            /* 1: waited at the queue head for memory    */
} vdq_serve_rec_t;
int vdq_serve(vdq_engine *e, const char *const *texts, const double *arrival_ms, int n,
              const vdq_serve_cfg_t *cfg, vdq_serve_rec_t *recs, vdq_result **outs);
/* Estimated final rows and peak HBM bytes of a query from the store
 * statistics, run alone through vdq_query (its joins followed by an
 * IndexSwap may be driven joins, whose run tables are counted).  Serving
 * admission uses the same estimate, without the driven joins' tables when
 * requests run batched. */
int vdq_estimate(vdq_engine *e, const char *text, double *rows, double *peak_bytes);

int vdq_result_shape(const vdq_result *r, int *ncols, uint64_t *nrows);
const char *vdq_result_header(const vdq_result *r, int col);
int vdq_result_column(const vdq_result *r, int col, uint32_t *host_out);
int vdq_result_stats(const vdq_result *r, vdq_stats_t *out);
int vdq_result_free(vdq_result *r);
/* Process-wide switches of the query executor, as the reference's static
 * QueryExecutor::ENABLE_* flags set from vdQuery's command line
 * (QueryExecutor.h:79-84, VedasQuery.cu:37,54,110-113):
 *   "bound_prescan"    (default 1) the join-variable bound pre-scan of
 *                      createVariableBound (QueryExecutor.cu:441-479);
 *   "bound_after_join" (default 1) the bound update after each join
 *                      (JoinQueryJob.cu:388-397; vdQuery's
 *                      -dis-bound-after-join clears it).
 * Both only narrow intermediate relations; results do not depend on them.
 * Set them between queries, not while another thread runs one. */
int vdq_set_option(const char *key, int value);
int vdq_get_option(const char *key, int *value);
/* Message of the calling thread's last failure. */
const char *vdq_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
