// Internal declarations shared by the translation units of libdyng_hip.so.
#pragma once

#include "../../include/dyng_hip.h"
#include "common.hpp"

#include <cstddef>
#include <memory>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dyng_hip {

constexpr int max_weights = 64;  ///< weight columns a graph may have on this backend

/// An error on its way to the C ABI (translated into a dyng_hip_status and the last message).
struct status_error : std::runtime_error {
  int code;
  uint32_t device_bits;
  status_error(int c, const std::string& m, uint32_t bits = 0)
      : std::runtime_error(m), code(c), device_bits(bits) {}
};

/// A batch that the apply refused before it changed anything (the batch semantics' error modes):
/// the graph and every result stay as they were.
struct batch_refused : status_error {
  using status_error::status_error;
};

[[noreturn]] void throw_hip(hipError_t e, const char* what, const char* file, int line);
[[noreturn]] inline void fail(int code, const std::string& m) { throw status_error(code, m); }
void set_last_error(const std::string& m);

#define DYNG_HIP_TRY(expr)                                                   \
  do {                                                                       \
    hipError_t dyng_e_ = (expr);                                             \
    if (dyng_e_ != hipSuccess) ::dyng_hip::throw_hip(dyng_e_, #expr, __FILE__, __LINE__); \
  } while (0)

/// Run `fn` and translate exceptions into a status (the C ABI never throws).
template <typename fn_t>
int guarded(fn_t&& fn) {
  try {
    fn();
    set_last_error("");
    return DYNG_HIP_OK;
  } catch (const status_error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("dyng: out of host memory");
    return DYNG_HIP_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return DYNG_HIP_RUNTIME_ERROR;
  }
}

/// Makes device `d` current for a scope and restores the caller's device afterwards (the
/// reference's scoped_device, cuda_runtime.hpp:77-94).
struct scoped_device {
  int previous = 0;
  explicit scoped_device(int d) {
    DYNG_HIP_TRY(hipGetDevice(&previous));
    if (previous != d) DYNG_HIP_TRY(hipSetDevice(d));
  }
  ~scoped_device() { (void)hipSetDevice(previous); }
  scoped_device(const scoped_device&) = delete;
  scoped_device& operator=(const scoped_device&) = delete;
};

/// A device array that grows on demand (hipMalloc; the contents are not kept on growth unless
/// `keep` is set).
template <typename T>
struct dbuf {
  T* p = nullptr;
  size_t n = 0;
  dbuf() = default;
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  ~dbuf() { release(); }
  void release() {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* reserve(size_t count, bool keep = false, hipStream_t stream = nullptr) {
    if (count <= n && p != nullptr) return p;
    const size_t want = count > 0 ? count : 1;
    T* q = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&q), want * sizeof(T));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw status_error(DYNG_HIP_OUT_OF_MEMORY,
                         "dyng: out of device memory (" + std::to_string(want * sizeof(T)) + " bytes)");
    }
    if (keep && p != nullptr && n > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(q, p, n * sizeof(T), hipMemcpyDeviceToDevice, stream));
      DYNG_HIP_TRY(hipStreamSynchronize(stream));
    }
    release();
    p = q;
    n = want;
    return p;
  }
  size_t bytes() const { return n * sizeof(T); }
};

/// Pinned host memory (hipHostMalloc).
template <typename T>
struct pinned {
  T* p = nullptr;
  pinned() {
    void* q = nullptr;
    if (hipHostMalloc(&q, sizeof(T), hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      throw status_error(DYNG_HIP_OUT_OF_MEMORY, "dyng: out of pinned host memory");
    }
    p = static_cast<T*>(q);
  }
  ~pinned() {
    if (p != nullptr) (void)hipHostFree(p);
  }
  pinned(const pinned&) = delete;
  pinned& operator=(const pinned&) = delete;
};

/// One direction of the device graph: per-vertex {offset, length} and capacity, and the slot
/// arena (neighbour ids, then num_weights weight columns of slot_cap entries each).
///
/// A row is two sorted runs in its slot range: the base [offset, offset + base) and the tail
/// [offset + base, offset + length). A batch deletes by marking entries dead (sign bit of the
/// neighbour id) and merges its new entries into the tail only; the row is rewritten as one
/// base when the tail or the dead entries outgrow their limits, or the row its capacity. The
/// row in the reference's order is the merge of the two runs by neighbour id, base entries first
/// among equal ones (they are older), without the dead entries.
struct row_store {
  dbuf<int2> meta;
  dbuf<int32_t> cap;
  dbuf<int32_t> base;  ///< per vertex: length of the base run
  dbuf<int32_t> dead;  ///< per vertex: dead entries in the base run
  dbuf<int32_t> nbr;
  dbuf<int32_t> w;
  int64_t slot_cap = 0;  ///< slots of the arena
  int64_t tail = 0;      ///< first slot of the free tail
  int64_t garbage = 0;   ///< slots left behind by relocated rows
};

/// The device-side counters of one apply, read back once at its end.
struct apply_control {
  unsigned long long inserted_edges, updated_edges, deleted_edges, ignored_deletions;
  unsigned long long dropped_self_loops, ignored_insertions;
  unsigned long long relocated_rows;
  int in_ops;
  int touched_out, touched_in;          ///< rows whose entries change
  int modes_out[4], modes_in[4];        ///< rows per merge mode (wave tail, workgroup tail, full, thread tail)
  int error_bits;
  long long tail_out, tail_in;        ///< arena tails (atomic bump allocation)
  long long stage_out, stage_in;      ///< staging slots used
  long long garbage_out, garbage_in;  ///< slots left behind by relocations
  // batch scan
  int min_id;
  int max_id;
  int self_loops_ins, self_loops_del;
  int min_weight;
  int pad0;
  // column summaries: target max per objective (set by the host), deltas (device)
  long long col_max[max_weights];
  long long col_dsum[max_weights];
  long long col_dcount[max_weights];
  int batch_wmax[max_weights];
  int batch_wmin[max_weights];
  long long col_bad[max_weights];  ///< weights below 1 per column (column summary)
  // merge plan: tail slots of the relocating rows, staging slots of the staged rows
  long long need_reloc_out, need_reloc_in, need_stage_out, need_stage_in;
  // error modes (on_missing_delete / on_existing_insert = error): the first offending operation
  // in the reference's order, as a rank (all ones: none); see k_check_pairs
  unsigned long long first_error;
  /// nonzero: the merges were queued without the host reading the plan, and the arenas could
  /// not take them as planned (or an error came first): the merges and any engine run queued
  /// behind them did nothing, and the host plans again (k_merge_gate)
  int hold;
  int pad1;
};

/// What the batch scan found (before anything changes) and where the batch is on the device.
struct batch_scan {
  int64_t n_after = 0;
  int64_t num_ins = 0, num_del = 0;
  const int32_t *ins_src = nullptr, *ins_dst = nullptr, *ins_w = nullptr, *del_src = nullptr, *del_dst = nullptr;
  std::vector<int64_t> wmax, wmin;  ///< per objective: largest / smallest insertion weight (min clamped to 1)
};

}  // namespace dyng_hip

struct dyng_hip_graph;
struct dyng_hip_mlp;
namespace dyng_hip {
void scan_batch(dyng_hip_graph* g, const dyng_hip_batch* b, hipStream_t s, batch_scan& sc);

/// One engine run in flight: engine_launch queues it on `s` (the checks, the launch and the
/// counters' read-back, nothing waited for), engine_finish waits for it and reports (or throws).
/// dyng_hip_sssp_update / _compute are the two back to back. Runs on one device are queued on one
/// stream, one after another: two persistent launches resident together gave wrong trees on
/// MI355X (the kernel keeps a little per-lane state in scratch memory), so runs never overlap.
struct engine_pending {
  dyng_hip_workspace* ws = nullptr;
  hipStream_t s = nullptr;
  bool launched = false;
  bool update = false;
  int bits = 0;
  int blocks = 0;
  int resident = 0;
  int64_t n = 0;
  int64_t timeout_us = 0;
  int64_t* distances = nullptr;
  bool held = false;  ///< set by engine_finish: the run did nothing (see engine_launch's `hold`)
  dyng_hip_counters counters{};
};
/// `hold`: queued behind apply_launch, the apply's
/// apply_hold flag: when the apply's merges were held back the run does nothing, and
/// engine_finish sets `held` (the caller launches it again after apply_finish).
engine_pending engine_launch(const dyng_hip_sssp_run* run, bool update, hipStream_t s, const int* hold = nullptr);
void engine_finish(engine_pending& pend, dyng_hip_counters* out);
/// The scheduler's pick for a scanned batch on G_t: features, BF16 inference, argmin (one sync).
int mlp_choose(dyng_hip_mlp* m, const dyng_hip_graph* g, const batch_scan& sc, const int32_t* parents, int64_t n_tree,
               int64_t last_invalidated, int64_t last_affected, hipStream_t s);
/// The same without waiting, on the network's own stream beside the apply: the prediction lands
/// in pinned host memory; the returned event marks it (and the end of the features' reads of the
/// rows). The apply waits for it before its merges and read-backs; mlp_pick then reads the choice.
hipEvent_t mlp_predict_async(dyng_hip_mlp* m, const dyng_hip_graph* g, const batch_scan& sc, const int32_t* parents,
                             int64_t n_tree, int64_t last_invalidated, int64_t last_affected);
int mlp_pick(const dyng_hip_mlp* m);
/// MOSP's combined graph of K trees (device parent arrays, a host array of K pointers), built on
/// the device as a solve-only graph (out-rows only) that a per-device cache owns and the next
/// call rebuilds; the edge count and weight sum come back with it. Callers hold pool().lock.
dyng_hip_graph* mosp_build_combined(int device, const int32_t* const* parents, int32_t num_objectives, int64_t n,
                                    int32_t source, int32_t base, const int32_t* terms, hipStream_t s,
                                    int64_t* edges, int64_t* weight_sum);
/// Grow the graph to n vertices without edges (new rows empty; no batch, no new version).
void graph_grow(dyng_hip_graph* g, int64_t n, hipStream_t s);
/// The combined graph of K trees (device pointer array `parents_dev`, device terms) as an edge
/// list into src / dst / w (capacity n * K); returns the edge count (and the weight sum).
int64_t mosp_combined_list(const int32_t* const* parents_dev, int32_t num_objectives, int64_t n, int32_t source,
                           int32_t base, const int32_t* terms_dev, int32_t* src, int32_t* dst, int32_t* w,
                           unsigned long long* totals, hipStream_t s, int64_t* weight_sum);
/// The batch from the combined graph of the old trees to that of the new ones (device pointer
/// arrays; capacity max(old_n, new_n) * K per list): deletions, then upserts.
void mosp_combined_diff(const int32_t* const* old_parents, int64_t old_n, const int32_t* const* new_parents,
                        int64_t new_n, int32_t num_objectives, int32_t source, int32_t base, const int32_t* terms_dev,
                        int32_t* del_src, int32_t* del_dst, int32_t* ins_src, int32_t* ins_dst, int32_t* ins_w,
                        unsigned long long* totals, hipStream_t s, int64_t* num_del, int64_t* num_ins);
/// The first `columns` objective costs of every vertex's path in the tree `parents` over the
/// graph's out-rows (device, n x columns, vertex-major); throws if a tree edge is missing.
/// Callers hold pool().lock (the doubling scratch is the per-device cache's).
void mosp_path_costs(const dyng_hip_graph* g, const int32_t* parents, int64_t n, int32_t source, int32_t columns,
                     int64_t* costs, hipStream_t s);
/// The apply after scan_batch. `wait` (optional): an event on another stream that must complete
/// before the merges change the rows (and before the apply's read-backs).
void apply_scanned(dyng_hip_graph* g, const batch_scan& sc, hipStream_t s, dyng_hip_apply_summary* out,
                   hipEvent_t wait = nullptr);
/// The apply in two halves: apply_launch queues the device work (the graph's vertex count is
/// n_after from then on; for a batch of up to kSpeculativeOps operations the merges are queued
/// without reading the merge plan back, guarded by a device flag, apply_hold) and the control
/// block's read-back; work queued on the stream after it sees the applied rows. apply_finish
/// waits for the stream, plans and merges again from the host if the guard held the merges
/// back (then `replanned` is set: engine runs queued in between did nothing), reports errors
/// and updates the graph's counts. Between the two only engine runs may be queued.
struct apply_pending {
  bool speculative = false;
  bool replanned = false;
  int64_t n_before = 0, N = 0;
  int vbits = 0;
  hipEvent_t wait = nullptr;
};
constexpr int64_t kSpeculativeOps = int64_t{1} << 17;
apply_pending apply_launch(dyng_hip_graph* g, const batch_scan& sc, hipStream_t s, dyng_hip_apply_summary* out,
                           hipEvent_t wait = nullptr);
void apply_finish(dyng_hip_graph* g, const batch_scan& sc, apply_pending& ap, hipStream_t s,
                  dyng_hip_apply_summary* out);
/// The device flag an engine run queued between apply_launch and apply_finish checks first.
const int* apply_hold(const dyng_hip_graph* g);
}  // namespace dyng_hip

/// The device graph behind the dyng_hip_graph handle.
struct dyng_hip_graph {
  int device = 0;
  dyng_hip_graph_props props{};
  int32_t K = 1;
  int64_t n = 0;      ///< vertices
  int64_t n_cap = 0;  ///< capacity of the per-vertex arrays
  int64_t m = 0;      ///< edges
  uint64_t version = 0;
  uint64_t state_id = 0;
  dyng_hip::row_store out, in;
  std::vector<int64_t> wmax, wsum, wmax_count;
  std::vector<int64_t> wmin;  ///< 1 when every weight of the column is >= 1, else 0 (or below)

  /// What the last apply found on G_t (before it changed anything), for the repair of a result
  /// one apply behind: per objective max(batch max, G_t max) and the smallest batch weight, and
  /// G_t's size and weight sums (the default delta); plus its summary.
  struct last_apply_info {
    bool valid = false;
    std::vector<int64_t> max_weight, batch_wmin, wsum;
    int64_t m = 0, n = 0, num_ins = 0;
    dyng_hip_apply_summary summary{};
  } last;

  // last apply: the effective operations (with holes, -1) and their weight-increase flags
  dyng_hip::dbuf<int32_t> del_src, del_dst, ins_src, ins_dst;
  dyng_hip::dbuf<uint8_t> ins_inc;
  dyng_hip::dbuf<int32_t> ins_fw;  ///< simple graphs: each insertion's pair weights after the batch (-1: gone)
  int64_t num_del_slots = 0, num_ins_slots = 0;

  // apply scratch
  dyng_hip::dbuf<unsigned long long> keys_a, keys_b;  ///< the apply's sort: out-row keys, then in-row keys
  dyng_hip::dbuf<uint32_t> vals_a, vals_b;
  dyng_hip::dbuf<unsigned char> sort_tmp, scan_tmp;
  dyng_hip::dbuf<int32_t> pend_alive, kill, pend_prefix, kill_prefix, pend_w;
  dyng_hip::dbuf<int32_t> ipend_alive, ikill, ipend_prefix, ikill_prefix, ipend_w;
  dyng_hip::dbuf<int32_t> iop_type, iop_wold, iop_wnew;
  dyng_hip::dbuf<int32_t> touched_out, touched_in, mark_out, mark_in;
  dyng_hip::dbuf<int32_t> row_a_out, row_b_out, row_a_in, row_b_in;  ///< per vertex: its operations
  dyng_hip::dbuf<int32_t> merge_lists;  ///< the merge plan's row lists, per direction and mode
  dyng_hip::dbuf<int32_t> stage_nbr, stage_w;
  int64_t stage_cap = 0;
  int mark_generation = 0;
  dyng_hip::dbuf<int32_t> b_ins_src, b_ins_dst, b_ins_w, b_del_src, b_del_dst;
  dyng_hip::dbuf<dyng_hip::apply_control> ctl;
  dyng_hip::pinned<dyng_hip::apply_control> host_ctl;
  /// events at the apply's stage boundaries (0-1 batch scan, 2-12 the update of the rows)
  hipEvent_t apply_ev[13] = {};
  bool apply_ev_recorded = false;

  dyng_hip_graph() = default;
  dyng_hip_graph(const dyng_hip_graph&) = delete;
  dyng_hip_graph& operator=(const dyng_hip_graph&) = delete;
  ~dyng_hip_graph() {
    for (hipEvent_t& e : apply_ev) {
      if (e != nullptr) (void)hipEventDestroy(e);
    }
  }
};
