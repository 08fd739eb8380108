/*
 * oracle.h -- CPU restatement of dynG's batch-dynamic SSSP path. TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker of the HIP product path and the CPU baseline of the benchmark. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * (dyng_amd/) never links, loads or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to the reference root).
 * Pinning (what checks each function beyond retyping it):
 *   orc_apply_batch      -- MOSP applyChangeBatch goldens: cpp/tests/data/mosp_graph_io/<case>/
 *                           applied/ (CSR, transposed CSR, weight-increase mask), 27 cases.
 *   orc_transpose        -- the same goldens (graphCsrTransposed*).
 *   orc_seq_* / orc_omp_* -- MOSP-OpenMP golden trees, cpp/tests/data/mosp_sssp (34 cases, every
 *                           objective: compute == init tree, update == updated tree, invalidated ==
 *                           stats.txt); hand cases of cpp/tests/algorithms/sssp/sssp_test.cpp;
 *                           randomized: compute == Dijkstra, update == compute on canonical trees,
 *                           invalidated/affected == brute force on the old/new trees.
 *   orc_dijkstra         -- brute force (Bellman-Ford with lowest-id ties) on small graphs and the
 *                           golden init trees.
 *   orc_check_tree       -- planted errors on the CPU (wrong distance, non-tight parent,
 *                           non-canonical parent).
 *   orc_build_from_edges -- the sssp_test.cpp hand graphs (parallel edges, duplicates merged).
 */
#ifndef DYNG_ORACLE_H
#define DYNG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* MOSP's DISTANCE_INF: INT64_MAX / 4 = 2^61 - 1 (cpp/include/dyng/core/types.hpp:91). */
#define ORC_INF INT64_C(2305843009213693951)

enum { ORC_OK = 0, ORC_INVALID = 1, ORC_NOT_SUPPORTED = 2, ORC_NOMEM = 3, ORC_INTERNAL = 4 };

/* Compact rows. Weights are objective-major: column k is w[k * m .. k * m + m). */
typedef struct orc_csr {
  int64_t n;
  int64_t m;
  int32_t K;
  int64_t* row_ptr; /* n + 1 */
  int32_t* col;     /* m */
  int32_t* w;       /* K * m */
} orc_csr;

/* graph_properties + batch_semantics (cpp/include/dyng/graph/graph_properties.hpp:67-221). */
typedef struct orc_props {
  int32_t directed;            /* 1 */
  int32_t sorted_rows;         /* row_order::sorted = 1, append = 0 */
  int32_t allow_parallel;      /* multi_edges::allow = 1, forbid = 0 */
  int32_t on_existing_insert;  /* 0 upsert, 1 error, 2 ignore */
  int32_t on_missing_delete;   /* 0 ignore, 1 error */
  int32_t on_self_loop;        /* 0 keep, 1 drop, 2 error */
  int32_t deletions_first;     /* 1 */
  int32_t allow_vertex_growth; /* 1 */
} orc_props;

/* apply_summary (cpp/include/dyng/graph/apply_summary.hpp:22-31). */
typedef struct orc_summary {
  int64_t inserted_edges, updated_edges, deleted_edges, ignored_deletions, dropped_self_loops;
  int64_t inserted_vertices, num_vertices_after, ignored_insertions;
} orc_summary;

/* The applied batch: the new rows plus apply_delta (cpp/src/graph/graph_impl.hpp:73-86). */
typedef struct orc_applied {
  orc_csr g;
  int64_t num_ins;
  int32_t* ins_src;
  int32_t* ins_dst;
  uint8_t* weight_increased; /* num_ins * K */
  int64_t num_del;
  int32_t* del_src;
  int32_t* del_dst;
  orc_summary summary;
} orc_applied;

/* One engine run (cpp/src/algorithms/sssp/problem.hpp:380-400, sssp_run). */
typedef struct orc_run {
  int64_t n;
  const int64_t* out_row;
  const int32_t* out_col;
  const int32_t* out_w;
  const int64_t* in_row; /* may be NULL for compute() of the OpenMP engine */
  const int32_t* in_col;
  const int32_t* in_w;
  const int32_t* changed_from;
  const int32_t* changed_to;
  int64_t num_changed;
  const int32_t* insert_heads;
  int64_t num_insert_heads;
  int32_t source;
  int32_t threads;
  int64_t delta;
  int64_t max_weight;
  int64_t* dist;   /* in: old tree (update); out: new tree */
  int32_t* parent; /* in: old tree (update); out: new tree */
  /* out: sssp_counters (problem.hpp:106-113) */
  int64_t invalidated, iterations, epochs, pushes, affected;
  int32_t packed_parents;
  /* out: timings of the four hooks in seconds (identify_affected, seed, loop, finalize) */
  double t_hooks[4];
} orc_run;

typedef struct orc_ws orc_ws; /* pooled workspace, reused across runs (ADR 0015) */

orc_ws* orc_ws_new(void);
void orc_ws_free(orc_ws* ws);

/* Errors: every function returns ORC_* and writes a message into err (errlen bytes). */
int orc_build_from_edges(int64_t n, int32_t K, int64_t count, const int32_t* src,
                         const int32_t* dst, const int32_t* weights, const orc_props* props,
                         orc_csr* out, char* err, int errlen);
int orc_apply_batch(const orc_csr* g, const orc_props* props, int64_t num_ins,
                    const int32_t* ins_src, const int32_t* ins_dst, const int32_t* ins_w,
                    int64_t num_del, const int32_t* del_src, const int32_t* del_dst,
                    int32_t threads, orc_applied* out, char* err, int errlen);
int orc_transpose(const orc_csr* g, orc_csr* out, int32_t threads);
void orc_csr_free(orc_csr* g);
void orc_applied_free(orc_applied* a);

/* Change lists of objective k (cpp/src/algorithms/sssp/sssp.cpp:563-576). The caller frees
 * *from and *to with orc_free. */
int64_t orc_build_changes(const orc_applied* a, int32_t k, int32_t** from, int32_t** to);
void orc_free(void* p);

int64_t orc_default_delta(int64_t m, int64_t n, int64_t weight_sum);
int32_t orc_distances_fit(int64_t n, int64_t max_weight);
int32_t orc_packs_parents(int64_t n, int64_t max_weight);
/* The device packing rule of MOSP-CUDA (kernels.cuh:82-92, cuda.cu:226-234): parent bits, or 0
 * for distance-only words. */
int32_t orc_device_parent_bits(int64_t n, int64_t max_weight);

int orc_seq_compute(orc_run* run, orc_ws* ws, char* err, int errlen);
int orc_seq_update(orc_run* run, orc_ws* ws, char* err, int errlen);
int orc_omp_compute(orc_run* run, orc_ws* ws, char* err, int errlen);
int orc_omp_update(orc_run* run, orc_ws* ws, char* err, int errlen);

void orc_dijkstra(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                  int32_t source, int64_t* dist, int32_t* parent);

/* check_sssp_tree (cpp/src/testing/check_sssp.cpp:33-125): counts[0..3] = distance mismatches,
 * inconsistent parents, non-canonical parents, parent mismatches (require_canonical). */
void orc_check_tree(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                    int32_t source, const int64_t* dist, const int32_t* parent,
                    int32_t require_canonical, int64_t counts[4]);

/* canonicalize_tree (cpp/src/algorithms/sssp/sssp.cpp:385-404). */
void orc_canonicalize(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                      int32_t source, const int64_t* dist, int32_t* parent);

#ifdef __cplusplus
}
#endif

#endif /* DYNG_ORACLE_H */
