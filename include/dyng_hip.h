/*
 * dyng_hip.h -- the C ABI of the MI355X-native sssp path (libdyng_hip.so).
 *
 * Plain pointers, sizes and POD structs; no C++ or torch types cross it. Two layers:
 *
 *  (1) Engine layer: what dynG's backend seam calls for the cuda backend today
 *      (cpp/src/algorithms/sssp/problem.hpp:557-571 sssp_cuda_update / sssp_cuda_compute with
 *      sssp_run, problem.hpp:380-400; the device graph upload of cpp/src/graph/graph.cpp:718-745
 *      and cpp/src/graph/device_graph.cu:55-150; the host batch apply of
 *      cpp/src/graph/apply_host.cpp:57-401, which this layer replaces with an in-place device
 *      apply). A C++ backend built on dynG's resources/framework calls these.
 *
 *  (2) API layer: what dynG's Python binding binds (python/bindings/sssp.cpp:77-135:
 *      native.sssp_compute / sssp_update / sssp_from_arrays and the result accessors; the graph
 *      of python/dyng/graph.py and the batch of python/dyng/batch.py). Handles are opaque; the
 *      host logic behind them (Step 0 checks, versions, poisoning, the change lists, the packing
 *      rule, delta) is the C++ of dyng_amd/csrc/api.cpp, which calls layer (1).
 *
 * Every function returns a dyng_hip_status; on failure dyng_hip_last_error() gives the message
 * (thread-local). The status values mirror dynG's exception classes (cpp/include/dyng/core/
 * error.hpp:28-164): invalid_argument_error, stale_result_error, not_supported_error,
 * out_of_memory_error, and a device error carrying the bits of util/device_error_flags.hpp.
 */
#ifndef DYNG_HIP_H
#define DYNG_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum dyng_hip_status {
  DYNG_HIP_OK = 0,
  DYNG_HIP_INVALID_ARGUMENT = 1, /* invalid_argument_error */
  DYNG_HIP_STALE_RESULT = 2,     /* stale_result_error */
  DYNG_HIP_NOT_SUPPORTED = 3,    /* not_supported_error */
  DYNG_HIP_OUT_OF_MEMORY = 4,    /* out_of_memory_error */
  DYNG_HIP_DEVICE_ERROR = 5,     /* a device error bit was raised (see device_errors) */
  DYNG_HIP_RUNTIME_ERROR = 6,    /* the HIP runtime reported an error (cuda_error) */
  DYNG_HIP_CAPACITY = 7          /* capacity_error: a storage cannot hold the batch */
} dyng_hip_status;

/* Device error bits (cpp/src/util/device_error_flags.hpp). */
#define DYNG_HIP_DEVICE_ERROR_CAPACITY 1u
#define DYNG_HIP_DEVICE_ERROR_INVALID_INPUT 2u
#define DYNG_HIP_DEVICE_ERROR_PARENT_CYCLE 4u
#define DYNG_HIP_DEVICE_ERROR_TIMEOUT 8u /* a grid barrier wait ran past tuning.barrier_timeout_us */
#define DYNG_HIP_DEVICE_ERROR_INTERNAL 0x80000000u

/* Engines (cpp/include/dyng/core/types.hpp, dyng::engine). */
#define DYNG_HIP_ENGINE_AUTOMATIC 0
#define DYNG_HIP_ENGINE_FUSED 1
#define DYNG_HIP_ENGINE_OPERATORS 2

/* The last error message of this thread ("" if none). */
const char* dyng_hip_last_error(void);

/* ============================================================================================
 * (1) Engine layer
 * ==========================================================================================*/

/* graph_properties + batch_semantics (cpp/include/dyng/graph/graph_properties.hpp:67-221). */
typedef struct dyng_hip_graph_props {
  int32_t directed;            /* 1: directed; 0: a batch edge (u,v) changes both arcs */
  int32_t sorted_rows;         /* row_order::sorted (1) or append (0); see the layout note */
  int32_t allow_parallel;      /* multi_edges::allow (1) or forbid (0) */
  int32_t on_existing_insert;  /* 0 upsert, 1 error, 2 ignore */
  int32_t on_missing_delete;   /* 0 ignore, 1 error (both error modes: checked before anything changes) */
  int32_t on_self_loop;        /* 0 keep, 1 drop (2 = error: checked on the host side) */
  int32_t deletions_first;     /* 1: deletions, then insertions */
  int32_t allow_vertex_growth; /* 1: an insertion naming v >= n grows the vertex set */
  double headroom;             /* spare slots per row: ceil(len * headroom) + 4 */
} dyng_hip_graph_props;

/* The device graph: out- and in-rows in slack storage (one slot range per row with spare
 * capacity, rows kept sorted by neighbour id, parallel edges in insertion order), so a batch
 * applies in place. Replaces the compact device copy plus in-edge build of device_graph.cu. */
typedef struct dyng_hip_graph dyng_hip_graph;

/* apply_summary (cpp/include/dyng/graph/apply_summary.hpp:22-41). */
typedef struct dyng_hip_apply_summary {
  int64_t inserted_edges, updated_edges, deleted_edges, ignored_deletions, dropped_self_loops;
  int64_t inserted_vertices, num_vertices_after, ignored_insertions, relocated_rows, compactions;
  int64_t host_plans; /* merge plans read back by the host before the merges (0: a batch of at most
                         2^17 operations whose merges the device planned and ran alone) */
  /* device time of the apply's stages (HIP events, ms; all 0 unless the environment sets
   * DYNG_HIP_STAGE_TIMES=1, since each timed event costs the stream a few microseconds): batch
   * scan, operation keys, key sort (both directions' keys in one sort), out-row pairing, (unused),
   * (unused), in-row pairing, the prefix scans of both directions (one scan), merge planning (with
   * its host round trip and any compaction), the in-place tail merges of both directions, the
   * whole-row merges through the staging arena */
  double stage_ms[11];
} dyng_hip_apply_summary;

/* A batch (cpp/include/dyng/graph/edge_batch.hpp:38-76): insertions (u, v, K weights,
 * insertion-major), deletions (u, v). `on_device` = 1: the arrays are device pointers. */
typedef struct dyng_hip_batch {
  const int32_t* ins_src;
  const int32_t* ins_dst;
  const int32_t* ins_w;
  int64_t num_ins;
  const int32_t* del_src;
  const int32_t* del_dst;
  int64_t num_del;
  int32_t num_weights;
  int32_t on_device;
} dyng_hip_batch;

/* Build the device graph from compact rows (row_ptr n+1, col m, w K*m objective-major), host
 * (from_device = 0) or device pointers. Replaces graph_access::prepare / build_device_graph
 * (cpp/src/graph/graph.cpp:1011-1018, device_graph.hpp:140-157). */
int dyng_hip_graph_create(int device, const dyng_hip_graph_props* props, int64_t n, int64_t m,
                          int32_t num_weights, const int64_t* row_ptr, const int32_t* col,
                          const int32_t* w, int32_t from_device, void* stream,
                          dyng_hip_graph** out);
/* Build the device graph from an edge list (src, dst, w count*K insertion-major), host or device
 * pointers, with the construction rules of build_from_edges_host
 * (cpp/src/graph/apply_host.cpp:608-690: repeated (u,v) of a simple graph merged, last wins). */
int dyng_hip_graph_create_from_edges(int device, const dyng_hip_graph_props* props, int64_t n,
                                     int64_t count, int32_t num_weights, const int32_t* src,
                                     const int32_t* dst, const int32_t* w, int32_t from_device,
                                     void* stream, dyng_hip_graph** out);
int dyng_hip_graph_destroy(dyng_hip_graph* g);
int dyng_hip_graph_num_vertices(const dyng_hip_graph* g, int64_t* n);
int dyng_hip_graph_num_edges(const dyng_hip_graph* g, int64_t* m);
/* Column summary of objective k (sssp.cpp:285-318 summarize_column, kept exact incrementally):
 * largest weight (>= 1) and weight sum. */
int dyng_hip_graph_column(const dyng_hip_graph* g, int32_t k, int64_t* max_weight,
                          int64_t* weight_sum);
/* Copy the compact out-rows (sorted by neighbour id) to host arrays sized n+1, m, K*m. */
int dyng_hip_graph_download(const dyng_hip_graph* g, int64_t* row_ptr, int32_t* col, int32_t* w);
/* Copy the compact in-rows (sorted by source id) to host arrays sized n+1, m, K*m. */
int dyng_hip_graph_download_in(const dyng_hip_graph* g, int64_t* row_ptr, int32_t* col,
                               int32_t* w);

/* Apply a batch in place on the device (MOSP applyChangeBatch semantics under the graph's
 * batch_semantics; apply_host.cpp:57-401 including the weight-increase classification of
 * :348-394). The effective deletions and insertions (after Step 0) and the per-objective
 * weight-increase flags stay on the device for the engine (dyng_hip_graph_changes). */
int dyng_hip_graph_apply(dyng_hip_graph* g, const dyng_hip_batch* batch, void* stream,
                         dyng_hip_apply_summary* summary);

/* The change list of the last apply, on the device (sssp.cpp:563-576 build_changes, consumed in
 * place: deletions, insertions and their weight-increase flags). Pointers stay valid until the
 * next apply. */
typedef struct dyng_hip_changes {
  const int32_t* del_src;
  const int32_t* del_dst;
  int64_t num_del;
  const int32_t* ins_src;
  const int32_t* ins_dst;
  const uint8_t* ins_increased; /* num_ins * num_weights flags */
  int64_t num_ins;
  int32_t num_weights;
} dyng_hip_changes;
int dyng_hip_graph_changes(const dyng_hip_graph* g, dyng_hip_changes* out);

/* Device scratch of the engine (sssp_cuda_workspace, problem.hpp:219-268): sized once for n
 * vertices, reused; stamps, in_far and flags are zero between runs. */
typedef struct dyng_hip_workspace dyng_hip_workspace;
int dyng_hip_workspace_create(int device, dyng_hip_workspace** out);
int dyng_hip_workspace_destroy(dyng_hip_workspace* ws);

/* Schedule knobs (the reference's only ones are delta and cuda_engine, sssp.hpp:78-94; they
 * change the schedule, never the result). 0 = default. */
#define DYNG_HIP_LAUNCH_PLAIN 0
#define DYNG_HIP_LAUNCH_GRAPH 1 /* the engine's launch built once as a hipGraph and replayed */

typedef struct dyng_hip_tuning {
  int32_t engine;          /* DYNG_HIP_ENGINE_* */
  int32_t grid_blocks;     /* workgroups of the persistent launch, at most what fits at once (0:
                              blocks_per_cu per CU) */
  int32_t blocks_per_cu;   /* workgroups per CU (0: all that fit for a solve from the sources, one
                              for a repair or when the library overlaps two runs on the device); a
                              caller that overlaps its own runs on one device asks for 1 */
  int32_t light_degree;    /* rows up to this length: edge-balanced inside a workgroup (0: 8192 in a repair, 4096 in a solve) */
  int32_t chunk_edges;     /* longer rows: split in chunks of this many edges, one wave each (0: 256) */
  int32_t launch;             /* DYNG_HIP_LAUNCH_*: a plain launch, or a hipGraph replayed per run */
  int32_t recompute;          /* update: recompute the tree from the source(s) instead of repairing it
                                 (the same tree; `invalidated` is then 0) */
  int32_t barrier_timeout_us; /* longest wait at a grid barrier before the run fails with
                                 DYNG_HIP_DEVICE_ERROR_TIMEOUT (0: 4 s; < 0: no wait at all, so
                                 the run fails at its first barrier: exercises the error path) */
} dyng_hip_tuning;

/* One engine run (sssp_run, problem.hpp:380-400), all arrays in device memory. */
typedef struct dyng_hip_sssp_run {
  const dyng_hip_graph* graph;
  int32_t objective;
  int32_t source;
  int64_t delta;      /* near-far width (> 0) */
  int64_t max_weight; /* largest weight of the objective before or after the batch (>= 1) */
  int64_t* distances; /* n; in: old tree (update); out: new tree */
  int32_t* parents;   /* n */
  int32_t check_cycles; /* update: 1 for a tree that did not come from this engine: its subtrees are
                           found by pointer jumping, which also detects a parent cycle (run even when
                           no tree edge changed); 0: the engine's own tree, walked top-down along its
                           tree edges */
  dyng_hip_workspace* workspace;
  dyng_hip_tuning tuning;
  /* several sources at once (no counterpart in the reference: the batched multi-instance path,
   * many independent trees of a disjoint union of graphs): source = -1 and these device ids;
   * null / 0 for the single source above */
  const int32_t* sources;
  int64_t num_sources;
  /* update: the token a previous run returned for the tree now in distances/parents (0: none).
   * When it is the workspace's last run's token, the workspace still holds the tree's packed
   * words and the run skips packing them (no counterpart in the reference, which packs the old
   * tree every time, fused.cuh:178-197). A caller that changes the arrays in between passes 0. */
  uint64_t tree_token;
} dyng_hip_sssp_run;

/* sssp_counters (problem.hpp:106-113) plus the relax kernel's work counters. */
typedef struct dyng_hip_counters {
  int64_t invalidated, iterations, epochs, pushes, affected;
  int32_t packed_parents;
  int32_t engine_used;
  uint32_t device_errors;
  int64_t relax_vertices; /* P: near-frontier vertices expanded */
  int64_t relax_edges;    /* E_s: out-edges scanned by the expansions */
  int64_t relax_improved; /* I: improving atomic minima */
  int64_t pull_edges;     /* in-edges scanned by the pull pass */
  int64_t jump_rounds;    /* pointer-jumping rounds run */
  int64_t jump_items;     /* vertices visited over all pointer-jumping rounds */
  int64_t pull_vertices;  /* candidates of the pull pass (invalidated + insertion heads) */
  double step1_ms;        /* device time of Step 1 (pack, invalidation, pull), s_memrealtime */
  double pull_ms;         /* of which the pull pass */
  double relax_ms;        /* device time of the near-far propagation (the relax iterations) */
  int64_t blocks;         /* workgroups of the persistent launch */
  double kernel_ms;       /* the persistent kernel's duration (HIP events on its stream) */
  double pack_ms;         /* Step 1 parts (s_memrealtime): packing the old tree and the roots, */
  double jump_ms;         /* the pointer-jumping rounds, */
  double unpack_ms;       /* and the final unpack */
  uint64_t tree_token;    /* the token of the tree this run left (0: its words are not kept) */
  int32_t words_kept;     /* 1: the run found the old tree's words in the workspace (no packing) */
  int32_t pad_;
} dyng_hip_counters;

/* sssp_cuda_update (cuda.cu:271-307): Step 1 + Step 2 on the graph after dyng_hip_graph_apply,
 * using its change list. One host synchronization (the control block read-back). */
int dyng_hip_sssp_update(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out);
/* sssp_cuda_compute (cuda.cu:309-335): the canonical tree from the source. */
int dyng_hip_sssp_compute(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out);

/* No counterpart in the reference: dyng_hip_sssp_update for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * same run and result, but the engine's scratch is one static device array of
 * DYNG_HIP_NOALLOC_SCRATCH_BYTES per device instead of run->workspace, which is ignored, as are
 * tree_token (the old tree is always packed from its arrays) and tuning.launch (a plain launch).
 * It calls no hipMalloc, hipFree or allocating library routine. A graph that needs more scratch
 * (88 B per vertex plus 2 B per row slot, plus 128 KiB) is refused with DYNG_HIP_INVALID_ARGUMENT
 * before anything runs. Every call on a device shares that device's array, which keeps the last
 * call's contents: calls that overlap on one device race. */
#define DYNG_HIP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int dyng_hip_sssp_update_noalloc(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out);

/* Device memory and synchronization for callers without their own HIP runtime plumbing (the
 * reference does this through rmm-like memory resources, cpp/include/dyng/core/memory.hpp):
 * device count, allocation, host<->device copies, a device-wide synchronize. */
int dyng_hip_device_count(int* count);
int dyng_hip_malloc(int device, int64_t bytes, void** out);
int dyng_hip_free(void* p);
int dyng_hip_memcpy_htod(void* dst, const void* src, int64_t bytes);
int dyng_hip_memcpy_dtoh(void* dst, const void* src, int64_t bytes);
int dyng_hip_synchronize(int device);

/* ============================================================================================
 * (2) API layer (python/bindings/sssp.cpp:77-135)
 * ==========================================================================================*/

typedef struct dyng_hip_options { /* sssp::options (sssp.hpp:78-94) */
  int64_t delta;
  int32_t objective;
  int32_t cuda_engine; /* DYNG_HIP_ENGINE_* */
  int32_t validate_inputs;
} dyng_hip_options;

typedef struct dyng_hip_stats { /* sssp::stats (sssp.hpp:100-117) + update_stats */
  int64_t affected, iterations, frontier_visits;
  int32_t fallback_used, converged, engine_used, packed_parents;
  dyng_hip_apply_summary batch;
  int64_t invalidated, epochs, pushes;
  dyng_hip_counters device; /* the engine run's counters */
  double apply_ms, engine_ms; /* stream time of the device apply and of the engine (events) */
} dyng_hip_stats;

typedef struct dyng_hip_result dyng_hip_result;

/* A graph handle with its version (graph::version(), graph.hpp). */
int dyng_hip_graph_version(const dyng_hip_graph* g, uint64_t* version);

/* native.sssp_compute (bindings/sssp.cpp:77-96 -> sssp::compute, sssp.hpp:415-421). */
int dyng_hip_api_compute(dyng_hip_graph* g, int32_t source, const dyng_hip_options* opt,
                         void* stream, dyng_hip_result** out);
/* Several trees in one run (no counterpart in the reference; the batched multi-instance path):
 * one result holding the trees of `count` distinct host sources, e.g. one per component of a
 * disjoint union of graphs. Its source is -1; updates repair all the trees in one launch. */
int dyng_hip_api_compute_multi(dyng_hip_graph* g, const int32_t* sources, int64_t count,
                               const dyng_hip_options* opt, void* stream, dyng_hip_result** out);
/* The repair half of sssp::update (sssp_problem::resume + the engine, sssp.cpp:691-823) for a
 * result exactly one dyng_hip_graph_apply behind the graph: with a clone of the result, several
 * schedules can be timed on one batch (the autotuner, the scheduler's training traces). */
int dyng_hip_api_repair(dyng_hip_graph* g, dyng_hip_result* r, void* stream, dyng_hip_stats* stats);
/* native.sssp_update (bindings/sssp.cpp:98-110 -> sssp::update, sssp.hpp:484-489). */
int dyng_hip_api_update(dyng_hip_graph* g, const dyng_hip_batch* batch, dyng_hip_result* r,
                        void* stream, dyng_hip_stats* stats);
/* dyng::update_each (cpp/include/dyng/update.hpp:311-364; Python dyng.update,
 * python/dyng/_update.py): the batch applied once and `count` results of the graph (distinct,
 * each current; any objectives and sources) repaired on the new graph, one stats entry each
 * (the apply time in the first). Every check runs before the batch is applied (nothing changes
 * on an error found there); after the commit a result whose repair fails is poisoned, the others
 * are still brought up to date, and the first error is returned. */
int dyng_hip_api_update_each(dyng_hip_graph* g, const dyng_hip_batch* batch, dyng_hip_result* const* results,
                             int64_t count, void* stream, dyng_hip_stats* stats);
/* native.sssp_from_arrays (bindings/sssp.cpp:112-135 -> result::from_arrays,
 * sssp.cpp:1118-1198): host arrays, imported, canonicalized and validated on the host. */
int dyng_hip_api_from_arrays(dyng_hip_graph* g, int32_t source, const int64_t* distances,
                             const int32_t* parents, int32_t canonicalize,
                             const dyng_hip_options* opt, void* stream, dyng_hip_result** out);
int dyng_hip_result_destroy(dyng_hip_result* r);
int dyng_hip_result_clone(const dyng_hip_result* r, void* stream, dyng_hip_result** out);
/* result accessors (sssp.hpp:183-234): device arrays of n entries, read-only (the next update
 * reuses the packed words the last run left for exactly these contents). */
int dyng_hip_result_arrays(const dyng_hip_result* r, const int64_t** distances,
                           const int32_t** parents, int64_t* n);
int dyng_hip_result_source(const dyng_hip_result* r, int32_t* source);
/* The work counters and timings of the engine run that left the result's tree (compute, update
 * or repair); no counterpart in the reference, whose compute() reports no stats. */
int dyng_hip_result_counters(const dyng_hip_result* r, dyng_hip_counters* out);
int dyng_hip_result_version(const dyng_hip_result* r, uint64_t* version);
int dyng_hip_result_options(const dyng_hip_result* r, dyng_hip_options* opt);
int dyng_hip_result_set_options(dyng_hip_result* r, const dyng_hip_options* opt);
/* Positions where two results' (distance, parent) pairs differ, counted on the device (0: the
 * trees are identical); no counterpart in the reference (it compares host vectors). */
int dyng_hip_result_equal(const dyng_hip_result* a, const dyng_hip_result* b, void* stream,
                          int64_t* differences);
/* Copy the arrays to host memory (to_vector). */
int dyng_hip_result_download(const dyng_hip_result* r, int64_t* distances, int32_t* parents);
/* Schedule knobs of this result's runs (not part of sssp::options). */
int dyng_hip_result_set_tuning(dyng_hip_result* r, const dyng_hip_tuning* tuning);

/* ============================================================================================
 * (3) The learned scheduler (no counterpart in the reference): batch features and a
 *     12 -> 64 -> 64 -> 5 MLP (ReLU) with FP32 master weights, trained on the device (FP32
 *     forward, backward, Adam) and evaluated in BF16. It predicts log2 of the time of five
 *     schedules of an update; every schedule gives the same tree.
 * ==========================================================================================*/

#define DYNG_HIP_MLP_INPUTS 12
#define DYNG_HIP_MLP_HIDDEN 64
#define DYNG_HIP_MLP_OUTPUTS 5
#define DYNG_HIP_MLP_TRAINED 5317 /* W1 64x12, b1, W2 64x64, b2, W3 5x64, b3 */
#define DYNG_HIP_MLP_PARAMS 5341  /* the trained ones, then x_mean[12], x_scale[12] */

/* What a batch looks like on G_t, before it is applied (the scheduler's raw features). */
typedef struct dyng_hip_batch_stats {
  int64_t num_ins, num_del;
  int64_t tree_deletions; /* deletions (u, v) with parent[v] == u in the given tree */
  int64_t touched;        /* operations whose tail vertex exists */
  double mean_log_degree; /* mean of log2(1 + out-degree) over those tails */
  int64_t max_degree;
  int64_t num_vertices, num_edges;
} dyng_hip_batch_stats;

/* `parents`: the tree's device parent array of n_tree entries (or null). */
int dyng_hip_batch_features(const dyng_hip_graph* g, const dyng_hip_batch* batch, const int32_t* parents,
                            int64_t n_tree, void* stream, dyng_hip_batch_stats* out);

typedef struct dyng_hip_mlp dyng_hip_mlp;
int dyng_hip_mlp_create(int device, uint64_t seed, dyng_hip_mlp** out);
int dyng_hip_mlp_destroy(dyng_hip_mlp* m);
/* All DYNG_HIP_MLP_PARAMS floats, host memory. */
int dyng_hip_mlp_get_params(const dyng_hip_mlp* m, float* params, int64_t count);
int dyng_hip_mlp_set_params(dyng_hip_mlp* m, const float* params, int64_t count);
/* FP32 forward + backward of the squared error over `count` samples (x: count x 12, t: count x 5,
 * host), summed over the known targets and divided by count x 5 (a NaN target, a schedule not
 * measured on that sample, adds nothing): the DYNG_HIP_MLP_TRAINED gradients and the loss. */
int dyng_hip_mlp_gradients(dyng_hip_mlp* m, const float* x, const float* t, int64_t count, float* grad,
                           float* loss);
/* Adam (lr, 0.9, 0.999, 1e-8) over `steps` minibatches of `batch` samples drawn without
 * replacement (reshuffled per pass); the final loss over all samples (FP32). */
int dyng_hip_mlp_train(dyng_hip_mlp* m, const float* x, const float* t, int64_t count, int32_t steps,
                       int32_t batch, float lr, uint64_t seed, float* final_loss);
/* BF16 inference: y (count x 5) for x (count x 12), host memory. */
int dyng_hip_mlp_predict(dyng_hip_mlp* m, const float* x, int64_t count, float* y);

/* sssp::update under the learned scheduler: the batch's features on G_t and the network's BF16
 * inference run on the device beside the apply, before it changes any row (mlp = null: the plain
 * schedule), then the update runs under the predicted fastest schedule if it is predicted at
 * least 15 % faster than the plain one, else under the plain one (*choice: its index).
 * last_invalidated and last_affected are the previous update's counters (the network's history
 * features). */
int dyng_hip_api_update_scheduled(dyng_hip_graph* g, const dyng_hip_batch* batch, dyng_hip_result* r,
                                  dyng_hip_mlp* mlp, int64_t last_invalidated, int64_t last_affected,
                                  void* stream, dyng_hip_stats* stats, int32_t* choice);

/* ============================================================================================
 * (4) MOSP's combined solve (cpp/src/algorithms/mosp/mosp.cpp:264-330, which runs it through
 *     sssp_solve_view, cpp/src/algorithms/sssp/sssp.cpp:1251-1286): the second caller of the
 *     engine, a static solve on a graph built from K shortest-path trees.
 * ==========================================================================================*/

typedef struct dyng_hip_mosp_stats {
  int64_t combined_edges;   /* edges of the combined graph */
  int64_t weight_sum;       /* their weight sum (the solve's default delta comes from it) */
  int64_t preference_scale; /* L = lcm(preferences), 1 without preferences */
  int64_t base;             /* L (K + 1): the largest combined weight */
  int64_t affected;         /* count_changes / an update: vertices whose combined (distance, parent) changed */
  int64_t deletions;        /* dyng_hip_mosp_update: combined edges the batch deleted */
  int64_t insertions;       /* and inserted or re-weighted */
  dyng_hip_counters solve;  /* the combined solve's (or repair's) engine counters */
} dyng_hip_mosp_stats;

/* The combined graph of the K trees (tree_parents: a host array of K device parent arrays of g's
 * n vertices; preferences: K values >= 1 with lcm <= 2^20, or null for all ones): for every
 * vertex v but the source and every objective k with p = parent_k[v] >= 0, the edge (p, v) once,
 * weighing L (K + 1) less L / pref[j] for every objective j whose tree holds it (MOSP's
 * combinedEdge, cpp/src/algorithms/mosp/openmp.cpp:39-118). Its canonical shortest-path tree
 * from the source goes to distances / parents (device, n); with count_changes they hold the
 * previous combined tree and `affected` counts the changed pairs. With cost_columns > 0,
 * path_costs (device, n x cost_columns, vertex-major) gets the first cost_columns weights of g
 * summed along each vertex's MOSP path, each tree edge (p, v) taken as the first (p, v) of p's
 * row (mosp_path_costs, cpp/src/algorithms/mosp/sequential.cpp:172-236); INF where the tree
 * does not reach. */
int dyng_hip_mosp_combined(const dyng_hip_graph* g, int32_t source, int32_t num_objectives,
                           const int32_t* const* tree_parents, const int32_t* preferences, int32_t count_changes,
                           int64_t* distances, int32_t* parents, int32_t cost_columns, int64_t* path_costs,
                           void* stream, dyng_hip_mosp_stats* out);

/* MOSP's combined state kept on the device between updates (the combined arrays of
 * mosp::result, python/dyng/mosp.py:186-403): the combined graph of the K trees kept as a device
 * graph and its canonical tree kept as an sssp result. dyng_hip_mosp_create builds both from the
 * trees (tree_parents: a host array of K device parent arrays of g's n vertices; preferences as
 * for dyng_hip_mosp_combined). dyng_hip_mosp_update takes the trees after an update of the same
 * graph (as many vertices or more) and turns the combined graph into theirs with one batch (the
 * combined edges the new trees lack deleted, the new or re-weighted ones upserted), applied and
 * repaired like any sssp update: the same tree the combined solve gives (the reference solves
 * it from scratch, mosp.cpp:319-330), `affected` counting the changed (distance, parent) pairs,
 * `solve` the repair's counters. Calls on one state are not thread-safe. */
typedef struct dyng_hip_mosp dyng_hip_mosp;
int dyng_hip_mosp_create(const dyng_hip_graph* g, int32_t source, int32_t num_objectives,
                         const int32_t* const* tree_parents, const int32_t* preferences, void* stream,
                         dyng_hip_mosp** out, dyng_hip_mosp_stats* stats);
int dyng_hip_mosp_update(dyng_hip_mosp* m, const dyng_hip_graph* g, const int32_t* const* tree_parents,
                         void* stream, dyng_hip_mosp_stats* stats);
/* The combined distances and the MOSP tree (device, n entries, read-only; valid until the next
 * update or the destroy). */
int dyng_hip_mosp_arrays(const dyng_hip_mosp* m, const int64_t** distances, const int32_t** parents,
                         int64_t* n);
/* The path costs of the state's MOSP tree over the first `columns` weights of g (device, n x
 * columns, vertex-major), as dyng_hip_mosp_combined computes them. */
int dyng_hip_mosp_path_costs(const dyng_hip_mosp* m, const dyng_hip_graph* g, int32_t columns,
                             int64_t* costs, void* stream);
int dyng_hip_mosp_destroy(dyng_hip_mosp* m);

/* ---- hyper_sssp's hypergraph (host side) ----
 * A weighted hypergraph with slack_csr storage (dyng::hypergraph<int32_t, int32_t>,
 * cpp/include/dyng/hypergraph/hypergraph.hpp; python/dyng/hypergraph.py), kept in host memory:
 * hyperedge i of the construction gets id i, members sorted and deduplicated. The device side of
 * hyper_sssp is a dyng_hip_graph of its line graph (vertex 0 a virtual source, vertex h + 1 the
 * hyperedge h), built from dyng_hip_hyper_line_graph and carried across each apply by
 * dyng_hip_hyper_line_graph_delta's edge batch. */
typedef struct dyng_hip_hypergraph dyng_hip_hypergraph;

/* hypergraph_properties' id_reuse and capacity_policy (hypergraph_properties.hpp:60-152). */
typedef struct {
  int32_t id_reuse; /* 0 erase_first (deletions, removals, additions, insertions), 1 insert_first */
  double initial_headroom;
  int64_t initial_slack;
  int64_t min_initial_payload;
  double grow_factor; /* 0: never grow (capacity_error) */
  int32_t compact_before_grow;
  int64_t max_payload; /* 0: 2^31 - 1 */
  int32_t row_alignment;
  int32_t compaction_headroom;
} dyng_hip_hyper_props;

/* hyperedge_batch_view (hyperedge_batch.hpp:41-51): hyperedge deletions, insertions (a CSR of
 * members, one weight each on a weighted hypergraph), incidence removals and additions. Host
 * arrays. */
typedef struct {
  const int64_t* delete_hyperedges;
  int64_t num_deletions;
  const int64_t* insert_offsets; /* num_insertions + 1 */
  const int32_t* insert_members;
  const int64_t* insert_weights;
  int64_t num_insertions;
  const int64_t* delete_incidence_hyperedge;
  const int32_t* delete_incidence_vertex;
  int64_t num_removals;
  const int64_t* insert_incidence_hyperedge;
  const int32_t* insert_incidence_vertex;
  int64_t num_additions;
} dyng_hip_hyper_batch;

/* hyper_apply_summary (hyper_apply_summary.hpp:20-36); the inserted ids go to a caller array. */
typedef struct {
  int64_t inserted_hyperedges, deleted_hyperedges, inserted_incidences, deleted_incidences;
  int64_t skipped_ops, ignored_deletions, dropped_members, num_hyperedges_after, id_bound_after;
} dyng_hip_hyper_summary;

/* The line graph's change of the last apply, for a result from `source`: the edges of its device
 * graph to delete and to insert or re-weight (upserts), and the unordered pairs the line graph
 * gained and lost (dyng::testing::line_graph_delta, hyper_sssp_reference.cpp:248-289). Host
 * arrays owned by the handle, valid until its next call. */
typedef struct {
  int64_t num_insertions;
  const int32_t* insert_src;
  const int32_t* insert_dst;
  const int32_t* insert_weights;
  int64_t num_deletions;
  const int32_t* delete_src;
  const int32_t* delete_dst;
  int64_t line_graph_insertions;
  int64_t line_graph_deletions;
} dyng_hip_hyper_delta;

/* hypergraph::from_lists / from_hyperedges (hypergraph.hpp): `weights` null for an unweighted
 * hypergraph (every weight >= 1 otherwise); members in [0, num_vertices). */
int dyng_hip_hypergraph_create(const int64_t* offsets, const int32_t* members, int64_t num_hyperedges,
                               const int64_t* weights, int32_t num_vertices, const dyng_hip_hyper_props* props,
                               dyng_hip_hypergraph** out);
int dyng_hip_hypergraph_destroy(dyng_hip_hypergraph* h);
/* hypergraph::clone (the ids, the store's state and the version kept). */
int dyng_hip_hypergraph_clone(const dyng_hip_hypergraph* h, dyng_hip_hypergraph** out);
/* num_vertices, num_hyperedges (live), hyperedge_id_bound, version, weighted. */
int dyng_hip_hypergraph_info(const dyng_hip_hypergraph* h, int32_t* num_vertices, int64_t* num_live,
                             int64_t* id_bound, int64_t* version, int32_t* weighted);
/* h2v(): the members of every id below the bound (offsets id_bound + 1 entries, a dead id's row
 * empty) and the weights by id (null when unweighted). Host arrays owned by the handle, valid
 * until its next call. */
int dyng_hip_hypergraph_rows(dyng_hip_hypergraph* h, const int64_t** offsets, const int32_t** members,
                             const int64_t** weights);
/* hypergraph::apply (hypergraph.cpp:573-1240): the phases, skip rules and counts of the reference,
 * the ids of its best-fit allocator under the hypergraph's id_reuse. Inserted weights above
 * `max_weight` (> 0) are refused before anything changes (hyper_sssp's validate_inputs);
 * `inserted_ids` gets num_insertions ids. A storage that cannot hold the batch under a policy
 * that does not grow raises DYNG_HIP_CAPACITY, the hypergraph unchanged. */
int dyng_hip_hypergraph_apply(dyng_hip_hypergraph* h, const dyng_hip_hyper_batch* b, int64_t max_weight,
                              dyng_hip_hyper_summary* out, int64_t* inserted_ids);
/* The device line graph of a hyper_sssp result from the vertex `source`, as an edge list: (0, h + 1)
 * of weight w(h) for the hyperedges that hold the source, (g + 1, h + 1) of weight w(h) for every
 * ordered pair of live hyperedges sharing a vertex (operators.cu:350-357,457: the head's node
 * weight). Host arrays owned by the handle, valid until its next call. */
int dyng_hip_hyper_line_graph(dyng_hip_hypergraph* h, int32_t source, int64_t* count, const int32_t** src,
                              const int32_t** dst, const int32_t** weight);
/* The change of that line graph over the last apply (see dyng_hip_hyper_delta). */
int dyng_hip_hyper_line_graph_delta(dyng_hip_hypergraph* h, int32_t source, dyng_hip_hyper_delta* out);

#ifdef __cplusplus
}
#endif

#endif /* DYNG_HIP_H */
