// The host side of the HIP sssp path: the semantics of dyng::sssp::compute / update /
// result::from_arrays (cpp/src/algorithms/sssp/sssp.cpp) around the device graph and the engine,
// exported as the API layer of include/dyng_hip.h (the calls python/bindings/sssp.cpp binds).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace dyng_hip {

namespace {
thread_local std::string last_error;
}

void set_last_error(const std::string& m) { last_error = m; }

void throw_hip(hipError_t e, const char* what, const char* file, int line) {
  (void)hipGetLastError();
  const int code = e == hipErrorOutOfMemory ? DYNG_HIP_OUT_OF_MEMORY : DYNG_HIP_RUNTIME_ERROR;
  throw status_error(code, std::string("dyng: HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) +
                               ") in " + what + " at " + file + ":" + std::to_string(line));
}

namespace {

constexpr int64_t kInf = sssp_infinity;

/// The pooled device workspaces of each device (ADR 0015: sized once, shared by the result runs
/// on the device; runs on one device are serialized by the pool's mutex). Slot 0 serves one
/// result at a time; an update of several results gives result i slot i % kWorkspaceSlots, so
/// those results keep their own packed words from one update to the next (up to kWorkspaceSlots
/// results) and kWorkspaceSlots of their runs can be queued ahead of the host.
constexpr int kWorkspaceSlots = 4;

struct workspace_pool {
  std::mutex lock;
  std::map<std::pair<int, int>, dyng_hip_workspace*> per_slot;
  dyng_hip_workspace* get(int device, int slot = 0) {
    auto it = per_slot.find({device, slot});
    if (it != per_slot.end()) return it->second;
    dyng_hip_workspace* ws = nullptr;
    if (dyng_hip_workspace_create(device, &ws) != DYNG_HIP_OK) fail(DYNG_HIP_OUT_OF_MEMORY, dyng_hip_last_error());
    per_slot[{device, slot}] = ws;
    return ws;
  }
};

workspace_pool& pool() {
  static workspace_pool* p = new workspace_pool();  // lives until the process ends
  return *p;
}

/// MOSP's defaultDelta (sssp.cpp:136-145).
int64_t default_delta(int64_t m, int64_t n, int64_t weight_sum) {
  if (m <= 0 || n <= 0) return 1;
  const double average_weight = static_cast<double>(weight_sum) / static_cast<double>(m);
  const double average_degree = static_cast<double>(m) / static_cast<double>(n);
  return std::max<int64_t>(1, static_cast<int64_t>(32.0 * average_weight / average_degree));
}

/// sssp_distances_fit (sssp.cpp:147-152).
bool distances_fit(int64_t n, int64_t max_weight) {
  const u64 weight = static_cast<u64>(std::max<int64_t>(max_weight, 1));
  const u64 hops = static_cast<u64>(std::max<int64_t>(n - 1, 1));
  return weight <= output_max_distance / hops;
}

void call(int status) {
  if (status != DYNG_HIP_OK) throw status_error(status, dyng_hip_last_error());
}

void expect_options(const dyng_hip_options& opt) {
  if (opt.delta < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: options.delta must be >= 0 (0 = automatic), got " + std::to_string(opt.delta));
  if (opt.objective < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: options.objective must be >= 0, got " + std::to_string(opt.objective));
  if (opt.cuda_engine < 0 || opt.cuda_engine > 2) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: options.cuda_engine is not an engine");
}

void expect_graph(const dyng_hip_graph* g, const dyng_hip_options& opt, const char* what) {
  if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, std::string(what) + ": null graph");
  if (opt.objective >= g->K) {
    fail(DYNG_HIP_INVALID_ARGUMENT, std::string(what) + ": options.objective " + std::to_string(opt.objective) +
                                        " is out of range for a graph with " + std::to_string(g->K) + " weight column(s)");
  }
}

}  // namespace

}  // namespace dyng_hip

/// The state behind a result handle (sssp_state, problem.hpp:275-312).
struct dyng_hip_result {
  int device = 0;
  int32_t source = 0;
  dyng_hip_options opt{};
  uint64_t version = 0;
  uint64_t graph_state = 0;
  bool poisoned = false;
  bool trusted = true;  ///< a tree this library computed, or an imported one that was validated
  /// the tree came from this engine (compute / update), so every tree edge is an edge of the graph
  bool engine_tree = true;
  int64_t n = 0;
  dyng_hip::dbuf<int64_t> distances;
  dyng_hip::dbuf<int32_t> parents;
  dyng_hip_tuning tuning{};
  /// several sources at once (source = -1): their ids on the device
  dyng_hip::dbuf<int32_t> sources;
  int64_t num_sources = 0;
  /// the token of the engine run that left this tree (its packed words may still be in the
  /// device's workspace); 0 for a tree the engine did not leave, or that changed since
  uint64_t tree_token = 0;
  /// the work counters of the engine run that left this tree (compute, update or repair)
  dyng_hip_counters last{};
};

using namespace dyng_hip;

namespace {

void expect_usable(const dyng_hip_result* r, const char* what) {
  if (r == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, std::string(what) + ": use of a moved-from result");
  if (r->poisoned) {
    throw status_error(DYNG_HIP_STALE_RESULT,
                       std::string("dyng: ") + what +
                           ": the result was left unusable by a failed update; recompute it with sssp::compute() "
                           "(or import a valid tree with sssp::result::from_arrays())");
  }
}

dyng_hip_sssp_run make_run(const dyng_hip_graph* g, dyng_hip_result* r, int64_t delta, int64_t max_weight) {
  dyng_hip_sssp_run run{};
  run.graph = g;
  run.objective = r->opt.objective;
  run.source = r->source;
  run.delta = delta;
  run.max_weight = max_weight;
  run.distances = r->distances.p;
  run.parents = r->parents.p;
  run.check_cycles = r->trusted && r->engine_tree ? 0 : 1;
  run.tree_token = r->tree_token;
  run.workspace = pool().get(g->device);
  run.tuning = r->tuning;
  if (r->opt.cuda_engine != DYNG_HIP_ENGINE_AUTOMATIC && run.tuning.engine == 0) run.tuning.engine = r->opt.cuda_engine;
  if (r->num_sources > 0) {
    run.source = -1;
    run.sources = r->sources.p;
    run.num_sources = r->num_sources;
  }
  return run;
}


/// Vertex growth: new vertices start unreachable (sssp.cpp:579-612).
void grow_result(dyng_hip_result* r, int64_t n, hipStream_t s) {
  if (n <= r->n) return;
  dbuf<int64_t> d;
  dbuf<int32_t> p;
  d.reserve(n);
  p.reserve(n);
  if (r->n > 0) {
    DYNG_HIP_TRY(hipMemcpyAsync(d.p, r->distances.p, r->n * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    DYNG_HIP_TRY(hipMemcpyAsync(p.p, r->parents.p, r->n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  }
  std::vector<int64_t> hd(static_cast<size_t>(n - r->n), kInf);
  std::vector<int32_t> hp(static_cast<size_t>(n - r->n), -1);
  DYNG_HIP_TRY(hipMemcpyAsync(d.p + r->n, hd.data(), hd.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
  DYNG_HIP_TRY(hipMemcpyAsync(p.p + r->n, hp.data(), hp.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  std::swap(r->distances.p, d.p);
  std::swap(r->distances.n, d.n);
  std::swap(r->parents.p, p.p);
  std::swap(r->parents.n, p.n);
  r->n = n;
}

void fill_stats(dyng_hip_stats* st, const dyng_hip_counters& c) {
  st->affected = c.affected;
  st->iterations = c.iterations;
  st->frontier_visits = c.pushes;
  st->fallback_used = 0;
  st->converged = 1;
  st->engine_used = c.engine_used;
  st->packed_parents = c.packed_parents;
  st->invalidated = c.invalidated;
  st->epochs = c.epochs;
  st->pushes = c.pushes;
  st->device = c;
}

/// MOSP's canonicalizeTree() (sssp.cpp:385-404) on host rows.
void canonicalize(int64_t n, const std::vector<int64_t>& row, const std::vector<int32_t>& col,
                  const int32_t* w, int32_t source, const std::vector<int64_t>& d, std::vector<int32_t>& p) {
  for (int64_t u = 0; u < n; ++u) {
    const int64_t du = d[u];
    if (du >= kInf / 2) continue;
    for (int64_t e = row[u]; e < row[u + 1]; ++e) {
      const int32_t v = col[e];
      if (v != source && u < p[v] && du + static_cast<int64_t>(w[e]) == d[v]) p[v] = static_cast<int32_t>(u);
    }
  }
}

/// validate_tree_sequential (sssp.cpp:408-456): the first violation in vertex order.
void validate_tree(int32_t source, const std::vector<int64_t>& d, const std::vector<int32_t>& p, int64_t bound) {
  const auto n = static_cast<int64_t>(d.size());
  if (d[source] != 0) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: the source " + std::to_string(source) + " must have distance 0, got " +
                                        std::to_string(d[source]));
  }
  if (p[source] != -1) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: the source " + std::to_string(source) + " must have parent -1, got " +
                                        std::to_string(p[source]));
  }
  for (int64_t v = 0; v < n; ++v) {
    const int64_t dv = d[v];
    const int32_t pv = p[v];
    if (dv >= kInf / 2) {
      if (pv != -1) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: unreachable vertex " + std::to_string(v) + " has parent " +
                                            std::to_string(pv) + " (expected -1)");
      }
      continue;
    }
    if (dv < 0 || dv > bound) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: the distance " + std::to_string(dv) + " of vertex " + std::to_string(v) +
                                          " is outside [0, (n - 1) * max weight = " + std::to_string(bound) + "]");
    }
    if (v == source) continue;
    if (pv < 0) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: reachable vertex " + std::to_string(v) + " (distance " +
                                          std::to_string(dv) + ") has no parent");
    }
    if (d[pv] >= kInf / 2) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: vertex " + std::to_string(v) + " has the unreachable parent " +
                                          std::to_string(pv));
    }
  }
  std::vector<char> state(static_cast<size_t>(n), 0);
  state[source] = 2;
  std::vector<int32_t> path;
  for (int64_t v = 0; v < n; ++v) {
    if (state[v] != 0 || d[v] >= kInf / 2) continue;
    path.clear();
    int32_t u = static_cast<int32_t>(v);
    while (state[u] == 0) {
      state[u] = 1;
      path.push_back(u);
      u = p[u];
    }
    if (state[u] != 2) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: the input shortest-path tree has a parent cycle through vertex " +
                                          std::to_string(u));
    }
    for (const int32_t x : path) state[x] = 2;
  }
}

struct event_pair {
  hipEvent_t a = nullptr, b = nullptr;
  event_pair() {
    DYNG_HIP_TRY(hipEventCreate(&a));
    DYNG_HIP_TRY(hipEventCreate(&b));
  }
  ~event_pair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  double ms() const {
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    return t;
  }
};

/// The engine part of an update (sssp_problem::resume + the engine run, sssp.cpp:691-823) for a
/// result that is one apply behind the graph; a failure poisons the result.
/// An apply queued ahead of a repair and not finished yet (apply_launch done, apply_finish not).
struct apply_in_flight {
  apply_pending* ap;
  const batch_scan* sc;
  dyng_hip_apply_summary* summary;
};

void repair_core(dyng_hip_graph* g, dyng_hip_result* r, int64_t delta, int64_t max_weight, hipStream_t s,
                 dyng_hip_counters& c, double& engine_ms, int choice = 0, const apply_in_flight* inflight = nullptr) {
  event_pair ev;
  try {
    grow_result(r, g->n, s);
    std::lock_guard<std::mutex> lk(pool().lock);
    auto run_for = [&]() {
      dyng_hip_sssp_run run = make_run(g, r, delta, max_weight);
      // the learned scheduler's schedules (they change the time, never the tree)
      switch (choice) {
        case 1: run.delta = 4 * delta; break;
        case 2: run.delta = std::max<int64_t>(1, delta / 4); break;
        case 3: run.tuning.blocks_per_cu = 2; break;  // the full grid (capped at what fits)
        case 4: run.tuning.recompute = 1; break;
        default: break;
      }
      return run;
    };
    dyng_hip_sssp_run run = run_for();
    DYNG_HIP_TRY(hipEventRecord(ev.a, s));
    r->tree_token = 0;
    // queued right behind the apply when it is in flight: one synchronization for both
    engine_pending pend = engine_launch(&run, true, s, inflight != nullptr ? apply_hold(g) : nullptr);
    DYNG_HIP_TRY(hipEventRecord(ev.b, s));
    if (inflight != nullptr) {
      try {
        apply_finish(g, *inflight->sc, *inflight->ap, s, inflight->summary);
      } catch (...) {
        try {
          engine_finish(pend, nullptr);
        } catch (...) {  // NOLINT(bugprone-empty-catch): the apply's error is the one reported
        }
        throw;
      }
    }
    engine_finish(pend, &c);
    if (pend.held) {
      // the apply's merges were planned again on the host (its arenas may have moved): the run now
      run = run_for();
      DYNG_HIP_TRY(hipEventRecord(ev.a, s));
      pend = engine_launch(&run, true, s);
      DYNG_HIP_TRY(hipEventRecord(ev.b, s));
      engine_finish(pend, &c);
    }
    r->tree_token = c.tree_token;
    r->last = c;
    DYNG_HIP_TRY(hipEventSynchronize(ev.b));
  } catch (...) {
    r->poisoned = true;
    throw;
  }
  r->version = g->version;
  r->trusted = true;
  r->engine_tree = true;
  engine_ms = ev.ms();
}

}  // namespace

extern "C" const char* dyng_hip_last_error(void) { return dyng_hip::last_error.c_str(); }

extern "C" int dyng_hip_api_compute(dyng_hip_graph* g, int32_t source, const dyng_hip_options* opt_in,
                                    void* stream, dyng_hip_result** out) {
  return guarded([&] {
    const dyng_hip_options opt = opt_in ? *opt_in : dyng_hip_options{0, 0, 0, 1};
    expect_options(opt);
    expect_graph(g, opt, "sssp::compute");
    const int64_t n = g->n;
    if (source < 0 || source >= n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: source " + std::to_string(source) + " is out of range [0, " +
                                          std::to_string(n) + ")");
    }
    const int k = opt.objective;
    if (g->wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: an edge has a weight below 1 in objective " + std::to_string(k) +
                                          "; sssp needs weights in [1, 2^31 - 1]");
    }
    if (!distances_fit(n, g->wmax[k])) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: distances up to " + std::to_string(g->wmax[k]) + " * " +
                                          std::to_string(n - 1) + " do not fit in 62 bits");
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto r = std::make_unique<dyng_hip_result>();
    r->device = g->device;
    r->source = source;
    r->opt = opt;
    r->version = g->version;
    r->graph_state = g->state_id;
    r->n = n;
    r->distances.reserve(n);
    r->parents.reserve(n);
    const int64_t delta = opt.delta > 0 ? opt.delta : default_delta(std::max<int64_t>(g->m, 1), n, g->wsum[k]);
    std::lock_guard<std::mutex> lk(pool().lock);
    dyng_hip_sssp_run run = make_run(g, r.get(), delta, g->wmax[k]);
    dyng_hip_counters c{};
    call(dyng_hip_sssp_compute(&run, s, &c));
    r->tree_token = c.tree_token;
    r->last = c;
    *out = r.release();
  });
}

extern "C" int dyng_hip_api_compute_multi(dyng_hip_graph* g, const int32_t* sources, int64_t count,
                                          const dyng_hip_options* opt_in, void* stream, dyng_hip_result** out) {
  return guarded([&] {
    if (out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: null output handle");
    const dyng_hip_options opt = opt_in ? *opt_in : dyng_hip_options{0, 0, 0, 1};
    expect_options(opt);
    expect_graph(g, opt, "sssp::compute");
    const int64_t n = g->n;
    if (count < 1 || count > n || sources == nullptr) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: between 1 and n sources are needed, got " + std::to_string(count));
    }
    std::vector<char> seen(static_cast<size_t>(n), 0);
    for (int64_t i = 0; i < count; ++i) {
      if (sources[i] < 0 || sources[i] >= n) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: source " + std::to_string(sources[i]) +
                                            " is out of range [0, " + std::to_string(n) + ")");
      }
      if (seen[sources[i]]++) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: source " + std::to_string(sources[i]) + " is listed twice");
    }
    const int k = opt.objective;
    if (g->wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: an edge has a weight below 1 in objective " + std::to_string(k) +
                                          "; sssp needs weights in [1, 2^31 - 1]");
    }
    if (!distances_fit(n, g->wmax[k])) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::compute: distances up to " + std::to_string(g->wmax[k]) + " * " +
                                          std::to_string(n - 1) + " do not fit in 62 bits");
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto r = std::make_unique<dyng_hip_result>();
    r->device = g->device;
    r->source = -1;
    r->opt = opt;
    r->version = g->version;
    r->graph_state = g->state_id;
    r->n = n;
    r->distances.reserve(n);
    r->parents.reserve(n);
    r->sources.reserve(count);
    r->num_sources = count;
    DYNG_HIP_TRY(hipMemcpyAsync(r->sources.p, sources, count * sizeof(int32_t), hipMemcpyHostToDevice, s));
    const int64_t delta = opt.delta > 0 ? opt.delta : default_delta(std::max<int64_t>(g->m, 1), n, g->wsum[k]);
    std::lock_guard<std::mutex> lk(pool().lock);
    dyng_hip_sssp_run run = make_run(g, r.get(), delta, g->wmax[k]);
    dyng_hip_counters c{};
    call(dyng_hip_sssp_compute(&run, s, &c));
    r->tree_token = c.tree_token;
    r->last = c;
    *out = r.release();
  });
}

extern "C" int dyng_hip_api_update(dyng_hip_graph* g, const dyng_hip_batch* batch, dyng_hip_result* r,
                                   void* stream, dyng_hip_stats* stats) {
  return guarded([&] {
    // ---- begin_update (sssp.cpp:620-640): nothing changes before these checks ----
    expect_usable(r, "sssp::update");
    if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: null graph");
    if (r->graph_state != g->state_id || r->version != g->version) {
      throw status_error(DYNG_HIP_STALE_RESULT,
                         "dyng: sssp::update: the result matches graph version " + std::to_string(r->version) +
                             " but the graph is at version " + std::to_string(g->version) +
                             " (or it belongs to another graph); recompute it");
    }
    if (r->device != g->device) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: the result is on another device");
    expect_graph(g, r->opt, "sssp::update");
    if (r->n != g->n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: the result has " + std::to_string(r->n) +
                                          " vertices but the graph has " + std::to_string(g->n));
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int k = r->opt.objective;
    // ---- prepare on G_t (sssp.cpp:643-689) ----
    batch_scan sc;
    scan_batch(g, batch, s, sc);
    if (batch->num_ins > 0 && sc.wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: an insertion has a weight below 1 in objective " +
                                          std::to_string(k) + "; sssp needs weights in [1, 2^31 - 1]");
    }
    const int64_t max_weight = std::max<int64_t>(std::max<int64_t>(sc.wmax[k], 1), g->wmax[k]);
    const int64_t delta =
        r->opt.delta > 0 ? r->opt.delta : default_delta(std::max<int64_t>(g->m, 1), g->n, g->wsum[k]);
    if (!distances_fit(sc.n_after, max_weight)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: distances up to " + std::to_string(max_weight) + " * " +
                                          std::to_string(sc.n_after - 1) + " do not fit in 62 bits");
    }
    // ---- commit: the device apply, with the engine queued right behind it (one synchronization
    // for both; apply_finish reports the apply's errors before the engine's) ----
    event_pair ev_apply;
    dyng_hip_apply_summary summary{};
    DYNG_HIP_TRY(hipEventRecord(ev_apply.a, s));
    apply_pending ap;
    try {
      ap = apply_launch(g, sc, s, &summary);
    } catch (const batch_refused&) {
      throw;  // refused before anything changed: the result stays current
    } catch (...) {
      r->poisoned = true;  // the graph may be partly updated
      throw;
    }
    DYNG_HIP_TRY(hipEventRecord(ev_apply.b, s));
    // ---- resume + the engine; a failure from here on poisons the result ----
    dyng_hip_counters c{};
    double engine_ms = 0.0;
    const apply_in_flight inflight{&ap, &sc, &summary};
    repair_core(g, r, delta, max_weight, s, c, engine_ms, 0, &inflight);
    if (stats != nullptr) {
      std::memset(stats, 0, sizeof(*stats));
      fill_stats(stats, c);
      stats->batch = summary;
      stats->apply_ms = ev_apply.ms();
      stats->engine_ms = engine_ms;
    }
  });
}

extern "C" int dyng_hip_api_update_each(dyng_hip_graph* g, const dyng_hip_batch* batch,
                                        dyng_hip_result* const* results, int64_t count, void* stream,
                                        dyng_hip_stats* stats) {
  return guarded([&] {
    // ---- every result's checks and preparation on G_t before the commit (update.hpp:263-364:
    // strong guarantee for what is found here) ----
    if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: null graph");
    if (results == nullptr || count < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: the list of results is empty (applying a batch "
                                      "without updating any result would leave every result of the graph "
                                      "stale; use the graph's apply)");
    }
    std::vector<const dyng_hip_result*> seen;
    seen.reserve(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) {
      const dyng_hip_result* r = results[i];
      if (r == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: result " + std::to_string(i) + " is null");
      if (std::find(seen.begin(), seen.end(), r) != seen.end()) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: result " + std::to_string(i) + " is listed twice");
      }
      seen.push_back(r);
      expect_usable(r, "dyng::update_each");
      if (r->graph_state != g->state_id || r->version != g->version) {
        throw status_error(DYNG_HIP_STALE_RESULT,
                           "dyng: dyng::update_each: result " + std::to_string(i) + " matches graph version " +
                               std::to_string(r->version) + " but the graph is at version " +
                               std::to_string(g->version) + " (or it belongs to another graph); recompute it");
      }
      if (r->device != g->device) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: result " + std::to_string(i) + " is on another device");
      }
      expect_graph(g, r->opt, "dyng::update_each");
      if (r->n != g->n) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: result " + std::to_string(i) + " has " +
                                            std::to_string(r->n) + " vertices but the graph has " +
                                            std::to_string(g->n));
      }
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    batch_scan sc;
    scan_batch(g, batch, s, sc);
    std::vector<int64_t> max_weight(static_cast<size_t>(count)), delta(static_cast<size_t>(count));
    for (int64_t i = 0; i < count; ++i) {
      const dyng_hip_result* r = results[i];
      const int k = r->opt.objective;
      if (batch->num_ins > 0 && sc.wmin[k] < 1) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: an insertion has a weight below 1 in objective " +
                                            std::to_string(k) + "; sssp needs weights in [1, 2^31 - 1]");
      }
      max_weight[i] = std::max<int64_t>(std::max<int64_t>(sc.wmax[k], 1), g->wmax[k]);
      delta[i] = r->opt.delta > 0 ? r->opt.delta : default_delta(std::max<int64_t>(g->m, 1), g->n, g->wsum[k]);
      if (!distances_fit(sc.n_after, max_weight[i])) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng::update_each: distances up to " + std::to_string(max_weight[i]) +
                                            " * " + std::to_string(sc.n_after - 1) + " do not fit in 62 bits");
      }
    }
    // ---- commit: the batch applied once ----
    event_pair ev_apply;
    dyng_hip_apply_summary summary{};
    DYNG_HIP_TRY(hipEventRecord(ev_apply.a, s));
    try {
      apply_scanned(g, sc, s, &summary);
    } catch (const batch_refused&) {
      throw;  // refused before anything changed: every result stays current
    } catch (...) {
      for (int64_t i = 0; i < count; ++i) results[i]->poisoned = true;  // the graph may be partly updated
      throw;
    }
    DYNG_HIP_TRY(hipEventRecord(ev_apply.b, s));
    // ---- every result repaired on G_{t+1}; a failing one is poisoned, the others still brought
    // up to date, and the first error reported ----
    std::exception_ptr first;
    if (count == 1) {
      dyng_hip_counters c{};
      double engine_ms = 0.0;
      try {
        repair_core(g, results[0], delta[0], max_weight[0], s, c, engine_ms);
        if (stats != nullptr) {
          std::memset(&stats[0], 0, sizeof(stats[0]));
          fill_stats(&stats[0], c);
          stats[0].batch = summary;
          stats[0].engine_ms = engine_ms;
        }
      } catch (...) {
        first = std::current_exception();
      }
    } else {
      // several results: result i runs on workspace slot i % kWorkspaceSlots, all of them queued
      // on the caller's stream, kWorkspaceSlots runs ahead of the host. They run one after
      // another: two persistent launches resident at once on MI355X gave wrong trees (each
      // lane keeps a little state in scratch memory, which the other launch's waves corrupted),
      // so runs never overlap
      for (int64_t i = 0; i < count; ++i) grow_result(results[i], g->n, s);
      std::lock_guard<std::mutex> lk(pool().lock);
      std::vector<engine_pending> pend(static_cast<size_t>(count));
      std::vector<char> ok(static_cast<size_t>(count), 0);
      auto finish = [&](int64_t i) {
        if (!ok[i]) return;
        ok[i] = 0;
        dyng_hip_result* r = results[i];
        dyng_hip_counters c{};
        try {
          engine_finish(pend[i], &c);
        } catch (...) {
          r->poisoned = true;
          if (!first) first = std::current_exception();
          return;
        }
        r->tree_token = c.tree_token;
        r->version = g->version;
        r->trusted = true;
        r->engine_tree = true;
        if (stats != nullptr) {
          std::memset(&stats[i], 0, sizeof(stats[i]));
          fill_stats(&stats[i], c);
          stats[i].batch = summary;
          stats[i].engine_ms = c.kernel_ms;
        }
      };
      for (int64_t i = 0; i < count; ++i) {
        // a workspace holds one run at a time: the run kWorkspaceSlots back used this slot
        if (i >= kWorkspaceSlots) finish(i - kWorkspaceSlots);
        dyng_hip_result* r = results[i];
        try {
          dyng_hip_sssp_run run = make_run(g, r, delta[i], max_weight[i]);
          run.workspace = pool().get(g->device, static_cast<int>(i % kWorkspaceSlots));
          r->tree_token = 0;
          pend[i] = engine_launch(&run, true, s);
          ok[i] = 1;
        } catch (...) {
          r->poisoned = true;
          if (!first) first = std::current_exception();
        }
      }
      for (int64_t i = std::max<int64_t>(0, count - kWorkspaceSlots); i < count; ++i) finish(i);
    }
    if (stats != nullptr) {
      DYNG_HIP_TRY(hipEventSynchronize(ev_apply.b));
      stats[0].apply_ms = ev_apply.ms();
    }
    if (first) std::rethrow_exception(first);
  });
}

extern "C" int dyng_hip_api_update_scheduled(dyng_hip_graph* g, const dyng_hip_batch* batch, dyng_hip_result* r,
                                             dyng_hip_mlp* mlp, int64_t last_invalidated, int64_t last_affected,
                                             void* stream, dyng_hip_stats* stats, int32_t* choice) {
  return guarded([&] {
    // ---- begin_update (sssp.cpp:620-640): nothing changes before these checks ----
    expect_usable(r, "sssp::update");
    if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: null graph");
    if (r->graph_state != g->state_id || r->version != g->version) {
      throw status_error(DYNG_HIP_STALE_RESULT,
                         "dyng: sssp::update: the result matches graph version " + std::to_string(r->version) +
                             " but the graph is at version " + std::to_string(g->version) +
                             " (or it belongs to another graph); recompute it");
    }
    if (r->device != g->device) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: the result is on another device");
    expect_graph(g, r->opt, "sssp::update");
    if (r->n != g->n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: the result has " + std::to_string(r->n) +
                                          " vertices but the graph has " + std::to_string(g->n));
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int k = r->opt.objective;
    // ---- prepare on G_t (sssp.cpp:643-689) ----
    batch_scan sc;
    scan_batch(g, batch, s, sc);
    if (batch->num_ins > 0 && sc.wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: an insertion has a weight below 1 in objective " +
                                          std::to_string(k) + "; sssp needs weights in [1, 2^31 - 1]");
    }
    const int64_t max_weight = std::max<int64_t>(std::max<int64_t>(sc.wmax[k], 1), g->wmax[k]);
    const int64_t delta =
        r->opt.delta > 0 ? r->opt.delta : default_delta(std::max<int64_t>(g->m, 1), g->n, g->wsum[k]);
    if (!distances_fit(sc.n_after, max_weight)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::update: distances up to " + std::to_string(max_weight) + " * " +
                                          std::to_string(sc.n_after - 1) + " do not fit in 62 bits");
    }
    // ---- the scheduler's features and inference, on G_t before anything changes; the apply's
    // own synchronization brings the prediction back ----
    hipEvent_t predicted = nullptr;
    if (mlp != nullptr) {
      predicted = mlp_predict_async(mlp, g, sc, r->parents.p, r->n, last_invalidated, last_affected);
    }
    // ---- commit: the device apply, the engine queued right behind it under the predicted
    // schedule (the prediction is waited for after the apply was queued) ----
    event_pair ev_apply;
    dyng_hip_apply_summary summary{};
    DYNG_HIP_TRY(hipEventRecord(ev_apply.a, s));
    apply_pending ap;
    try {
      ap = apply_launch(g, sc, s, &summary, predicted);
    } catch (const batch_refused&) {
      throw;  // refused before anything changed: the result stays current
    } catch (...) {
      r->poisoned = true;  // the graph may be partly updated
      throw;
    }
    DYNG_HIP_TRY(hipEventRecord(ev_apply.b, s));
    // ---- resume + the engine; a failure from here on poisons the result ----
    dyng_hip_counters c{};
    double engine_ms = 0.0;
    if (predicted != nullptr) DYNG_HIP_TRY(hipEventSynchronize(predicted));
    const int pick = mlp != nullptr ? mlp_pick(mlp) : 0;
    if (choice != nullptr) *choice = pick;
    const apply_in_flight inflight{&ap, &sc, &summary};
    repair_core(g, r, delta, max_weight, s, c, engine_ms, pick, &inflight);
    if (stats != nullptr) {
      std::memset(stats, 0, sizeof(*stats));
      fill_stats(stats, c);
      stats->batch = summary;
      stats->apply_ms = ev_apply.ms();
      stats->engine_ms = engine_ms;
    }
  });
}

extern "C" int dyng_hip_api_repair(dyng_hip_graph* g, dyng_hip_result* r, void* stream, dyng_hip_stats* stats) {
  return guarded([&] {
    expect_usable(r, "sssp::repair");
    if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::repair: null graph");
    if (r->graph_state != g->state_id || r->version + 1 != g->version || !g->last.valid) {
      throw status_error(DYNG_HIP_STALE_RESULT,
                         "dyng: sssp::repair: the result matches graph version " + std::to_string(r->version) +
                             " but repair needs a result exactly one apply behind the graph (version " +
                             std::to_string(g->version) + ")");
    }
    if (r->device != g->device) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::repair: the result is on another device");
    expect_graph(g, r->opt, "sssp::repair");
    if (r->n > g->n) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::repair: the result has more vertices than the graph");
    const int k = r->opt.objective;
    if (g->last.num_ins > 0 && g->last.batch_wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::repair: the last batch inserted a weight below 1 in objective " +
                                          std::to_string(k) + "; sssp needs weights in [1, 2^31 - 1]");
    }
    const int64_t max_weight = g->last.max_weight[k];
    const int64_t delta =
        r->opt.delta > 0 ? r->opt.delta : default_delta(std::max<int64_t>(g->last.m, 1), g->last.n, g->last.wsum[k]);
    if (!distances_fit(g->n, max_weight)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::repair: distances up to " + std::to_string(max_weight) + " * " +
                                          std::to_string(g->n - 1) + " do not fit in 62 bits");
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    dyng_hip_counters c{};
    double engine_ms = 0.0;
    repair_core(g, r, delta, max_weight, s, c, engine_ms);
    if (stats != nullptr) {
      std::memset(stats, 0, sizeof(*stats));
      fill_stats(stats, c);
      stats->batch = g->last.summary;
      stats->engine_ms = engine_ms;
    }
  });
}

extern "C" int dyng_hip_api_from_arrays(dyng_hip_graph* g, int32_t source, const int64_t* distances,
                                        const int32_t* parents, int32_t canonicalize_tree, const dyng_hip_options* opt_in,
                                        void* stream, dyng_hip_result** out) {
  return guarded([&] {
    const dyng_hip_options opt = opt_in ? *opt_in : dyng_hip_options{0, 0, 0, 1};
    expect_options(opt);
    expect_graph(g, opt, "sssp::result::from_arrays");
    const int64_t n = g->n;
    if (source < 0 || source >= n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::from_arrays: source " + std::to_string(source) +
                                          " is out of range [0, " + std::to_string(n) + ")");
    }
    const int k = opt.objective;
    if (g->wmin[k] < 1) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::from_arrays: an edge has a weight below 1 in objective " +
                                          std::to_string(k));
    }
    // import (normalize_imported_tree, sssp.cpp:533-558)
    std::vector<int64_t> d(distances, distances + n);
    std::vector<int32_t> p(parents, parents + n);
    for (int64_t v = 0; v < n; ++v) {
      if (d[v] >= kInf / 2) d[v] = kInf;
      if (p[v] < -1 || p[v] >= n) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::from_arrays: vertex " + std::to_string(v) + " has parent " +
                                            std::to_string(p[v]) + ", outside [-1, " + std::to_string(n) + ")");
      }
    }
    if (canonicalize_tree) {
      std::vector<int64_t> row(static_cast<size_t>(n) + 1);
      std::vector<int32_t> col(static_cast<size_t>(std::max<int64_t>(g->m, 1)));
      std::vector<int32_t> w(static_cast<size_t>(std::max<int64_t>(g->m, 1)) * g->K);
      call(dyng_hip_graph_download(g, row.data(), col.data(), w.data()));
      canonicalize(n, row, col, w.data() + static_cast<int64_t>(k) * g->m, source, d, p);
    }
    if (opt.validate_inputs) {
      if (!distances_fit(n, g->wmax[k])) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::from_arrays: distances up to " + std::to_string(g->wmax[k]) +
                                            " * " + std::to_string(n - 1) + " do not fit in 62 bits");
      }
      validate_tree(source, d, p, g->wmax[k] * std::max<int64_t>(n - 1, 1));
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto r = std::make_unique<dyng_hip_result>();
    r->device = g->device;
    r->source = source;
    r->opt = opt;
    r->version = g->version;
    r->graph_state = g->state_id;
    r->trusted = opt.validate_inputs != 0;
    r->engine_tree = false;
    r->n = n;
    r->distances.reserve(n);
    r->parents.reserve(n);
    if (n > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(r->distances.p, d.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(r->parents.p, p.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    *out = r.release();
  });
}

namespace {

/// Count the positions where two arrays of the same length differ (one atomic per wave).
template <typename T>
__global__ void k_count_differences(const T* a, const T* b, int64_t n, unsigned long long* out) {
  unsigned long long diff = 0;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < n; i += int64_t{gridDim.x} * blockDim.x) {
    diff += a[i] != b[i] ? 1 : 0;
  }
  diff = wave_sum(diff);
  if (lane() == 0 && diff) atomicAdd(out, diff);
}

}  // namespace

extern "C" int dyng_hip_result_equal(const dyng_hip_result* a, const dyng_hip_result* b, void* stream,
                                     int64_t* differences) {
  return guarded([&] {
    expect_usable(a, "sssp::result::equal");
    expect_usable(b, "sssp::result::equal");
    if (differences == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::equal: null output");
    if (a->device != b->device) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::equal: the results are on other devices");
    if (a->n != b->n) {
      *differences = std::max(a->n, b->n);
      return;
    }
    scoped_device guard(a->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    dbuf<unsigned long long> count;
    count.reserve(1);
    DYNG_HIP_TRY(hipMemsetAsync(count.p, 0, sizeof(unsigned long long), s));
    if (a->n > 0) {
      const unsigned grid = static_cast<unsigned>(std::min<int64_t>((a->n + 255) / 256, 4096));
      hipLaunchKernelGGL(k_count_differences<int64_t>, dim3(grid), dim3(256), 0, s, a->distances.p, b->distances.p,
                         a->n, count.p);
      hipLaunchKernelGGL(k_count_differences<int32_t>, dim3(grid), dim3(256), 0, s, a->parents.p, b->parents.p, a->n,
                         count.p);
      DYNG_HIP_TRY(hipGetLastError());
    }
    unsigned long long h = 0;
    DYNG_HIP_TRY(hipMemcpyAsync(&h, count.p, sizeof(h), hipMemcpyDeviceToHost, s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    *differences = static_cast<int64_t>(h);
  });
}

extern "C" int dyng_hip_result_destroy(dyng_hip_result* r) {
  return guarded([&] { delete r; });
}

extern "C" int dyng_hip_result_clone(const dyng_hip_result* r, void* stream, dyng_hip_result** out) {
  return guarded([&] {
    expect_usable(r, "sssp::result::clone");
    scoped_device guard(r->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto c = std::make_unique<dyng_hip_result>();
    c->device = r->device;
    c->source = r->source;
    c->opt = r->opt;
    c->version = r->version;
    c->graph_state = r->graph_state;
    c->trusted = r->trusted;
    c->engine_tree = r->engine_tree;
    c->tuning = r->tuning;
    c->n = r->n;
    c->distances.reserve(r->n);
    c->parents.reserve(r->n);
    if (r->num_sources > 0) {
      c->sources.reserve(r->num_sources);
      c->num_sources = r->num_sources;
      DYNG_HIP_TRY(hipMemcpyAsync(c->sources.p, r->sources.p, r->num_sources * sizeof(int32_t),
                                  hipMemcpyDeviceToDevice, s));
    }
    if (r->n > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(c->distances.p, r->distances.p, r->n * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(c->parents.p, r->parents.p, r->n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    *out = c.release();
  });
}

extern "C" int dyng_hip_result_arrays(const dyng_hip_result* r, const int64_t** distances, const int32_t** parents,
                                      int64_t* n) {
  return guarded([&] {
    expect_usable(r, "sssp::result::distances");
    *distances = r->distances.p;
    *parents = r->parents.p;
    *n = r->n;
  });
}

extern "C" int dyng_hip_result_source(const dyng_hip_result* r, int32_t* source) {
  return guarded([&] { *source = r ? r->source : -1; });
}

extern "C" int dyng_hip_result_counters(const dyng_hip_result* r, dyng_hip_counters* out) {
  return guarded([&] {
    if (r == nullptr || out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: null result or output");
    *out = r->last;
  });
}

extern "C" int dyng_hip_result_version(const dyng_hip_result* r, uint64_t* version) {
  return guarded([&] { *version = r ? r->version : 0; });
}

extern "C" int dyng_hip_result_options(const dyng_hip_result* r, dyng_hip_options* opt) {
  return guarded([&] {
    expect_usable(r, "sssp::result::get_options");
    *opt = r->opt;
  });
}

extern "C" int dyng_hip_result_set_options(dyng_hip_result* r, const dyng_hip_options* opt) {
  return guarded([&] {
    expect_usable(r, "sssp::result::set_options");
    expect_options(*opt);
    if (opt->objective != r->opt.objective) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp::result::set_options: the objective is fixed at compute() (" +
                                          std::to_string(r->opt.objective) + "); got " + std::to_string(opt->objective));
    }
    r->opt = *opt;
  });
}

extern "C" int dyng_hip_result_download(const dyng_hip_result* r, int64_t* distances, int32_t* parents) {
  return guarded([&] {
    expect_usable(r, "sssp::result::to_vector");
    scoped_device guard(r->device);
    if (r->n > 0) {
      DYNG_HIP_TRY(hipMemcpy(distances, r->distances.p, r->n * sizeof(int64_t), hipMemcpyDeviceToHost));
      DYNG_HIP_TRY(hipMemcpy(parents, r->parents.p, r->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
  });
}

extern "C" int dyng_hip_result_set_tuning(dyng_hip_result* r, const dyng_hip_tuning* tuning) {
  return guarded([&] {
    if (r == nullptr || tuning == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: null argument");
    r->tuning = *tuning;
  });
}

extern "C" int dyng_hip_device_count(int* count) {
  return guarded([&] {
    int c = 0;
    const hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      c = 0;
    }
    *count = c;
  });
}

extern "C" int dyng_hip_malloc(int device, int64_t bytes, void** out) {
  return guarded([&] {
    scoped_device guard(device);
    void* p = nullptr;
    if (bytes > 0) {
      const hipError_t e = hipMalloc(&p, static_cast<size_t>(bytes));
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(DYNG_HIP_OUT_OF_MEMORY, "dyng: out of device memory (" + std::to_string(bytes) + " bytes)");
      }
    }
    *out = p;
  });
}

extern "C" int dyng_hip_free(void* p) {
  return guarded([&] {
    if (p != nullptr) DYNG_HIP_TRY(hipFree(p));
  });
}

extern "C" int dyng_hip_memcpy_htod(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes > 0) DYNG_HIP_TRY(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
  });
}

extern "C" int dyng_hip_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
  return guarded([&] {
    if (bytes > 0) DYNG_HIP_TRY(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
  });
}

extern "C" int dyng_hip_synchronize(int device) {
  return guarded([&] {
    scoped_device guard(device);
    DYNG_HIP_TRY(hipDeviceSynchronize());
  });
}

// ---------------------------------------------------------------------------------------------
// MOSP's combined solve (cpp/src/algorithms/mosp/mosp.cpp:264-330)
// ---------------------------------------------------------------------------------------------

namespace {

/// MOSP's checks of the trees and the preferences, and the combined weights' terms
/// (preferenceScale, mosp.cpp:114-131: L = lcm(preferences) <= 2^20; terms L / pref[k]; base
/// L (K + 1)).
struct mosp_terms {
  int64_t scale = 1;
  int64_t base = 0;
  std::vector<int32_t> terms;
};

mosp_terms mosp_check(const dyng_hip_graph* g, int32_t source, int32_t num_objectives,
                      const int32_t* const* tree_parents, const int32_t* preferences) {
  if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: null graph");
  const int64_t n = g->n;
  if (num_objectives < 1 || num_objectives > 64) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: 1 to 64 objectives, got " + std::to_string(num_objectives));
  }
  if (tree_parents == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the trees are required");
  for (int k = 0; k < num_objectives; ++k) {
    if (tree_parents[k] == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: tree " + std::to_string(k) + " is null");
  }
  if (n > 0 && (source < 0 || source >= n)) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: source " + std::to_string(source) + " is out of range [0, " +
                                        std::to_string(n) + ")");
  }
  mosp_terms t;
  if (preferences != nullptr) {
    for (int k = 0; k < num_objectives; ++k) {
      const int64_t pref = preferences[k];
      if (pref < 1) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: preferences[" + std::to_string(k) + "] is " +
                                                        std::to_string(pref) + "; a preference must be >= 1");
      int64_t a = t.scale, b = pref;
      while (b != 0) {
        const int64_t r = a % b;
        a = b;
        b = r;
      }
      t.scale = t.scale / a * pref;
      if (t.scale > (int64_t{1} << 20)) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: lcm(preferences) exceeds 2^20");
    }
  }
  t.terms.resize(static_cast<size_t>(num_objectives));
  for (int k = 0; k < num_objectives; ++k) {
    t.terms[k] = static_cast<int32_t>(preferences != nullptr ? t.scale / preferences[k] : t.scale);
  }
  t.base = t.scale * (num_objectives + 1);
  return t;
}

}  // namespace

extern "C" int dyng_hip_mosp_combined(const dyng_hip_graph* g, int32_t source, int32_t num_objectives,
                                      const int32_t* const* tree_parents, const int32_t* preferences,
                                      int32_t count_changes, int64_t* distances, int32_t* parents,
                                      int32_t cost_columns, int64_t* path_costs, void* stream,
                                      dyng_hip_mosp_stats* out) {
  return guarded([&] {
    const mosp_terms mt = mosp_check(g, source, num_objectives, tree_parents, preferences);
    const int64_t n = g->n;
    if (distances == nullptr || parents == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the output arrays are required");
    if (cost_columns < 0 || cost_columns > g->K || (cost_columns > 0 && path_costs == nullptr)) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: cost_columns must be in [0, " + std::to_string(g->K) +
                                          "] with an output array");
    }
    const int64_t scale = mt.scale;
    const std::vector<int32_t>& terms = mt.terms;
    const int64_t base = mt.base;
    dyng_hip_mosp_stats st{};
    st.preference_scale = scale;
    st.base = base;
    if (n == 0) {
      if (out != nullptr) *out = st;
      return;
    }
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    // one lock over the build, the solve and the costs: the combined graph and its scratch are
    // a per-device cache that the next call reuses, and the solve takes the pooled workspace
    std::lock_guard<std::mutex> lk(pool().lock);
    int64_t edges = 0, weight_sum = 0;
    dyng_hip_graph* cg = mosp_build_combined(g->device, tree_parents, num_objectives, n, source,
                                             static_cast<int32_t>(base), terms.data(), s, &edges, &weight_sum);
    st.combined_edges = edges;
    st.weight_sum = weight_sum;
    // the static solve of sssp_solve_view: the automatic width of the combined graph, the base
    // weight as the largest weight; with count_changes a recompute that counts the pairs it
    // changes against the arrays' previous tree
    dyng_hip_sssp_run run{};
    run.graph = cg;
    run.objective = 0;
    run.source = source;
    run.delta = default_delta(std::max<int64_t>(edges, 1), n, weight_sum);
    run.max_weight = base;
    run.distances = distances;
    run.parents = parents;
    run.workspace = pool().get(g->device);
    dyng_hip_counters c{};
    if (count_changes) {
      run.tuning.recompute = 1;
      call(dyng_hip_sssp_update(&run, s, &c));
      st.affected = c.affected;
    } else {
      call(dyng_hip_sssp_compute(&run, s, &c));
    }
    st.solve = c;
    if (cost_columns > 0) mosp_path_costs(g, parents, n, source, cost_columns, path_costs, s);
    // the cached graph must not be rebuilt while this call's work is queued
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    if (out != nullptr) *out = st;
  });
}

// ---- MOSP's combined state kept on the device: the combined graph and its tree ----

struct dyng_hip_mosp {
  int device = 0;
  int32_t source = 0;
  int32_t K = 0;
  mosp_terms terms;
  dyng_hip_graph* cg = nullptr;   ///< the combined graph (a graph of its own, one weight column)
  dyng_hip_result* cr = nullptr;  ///< its canonical tree from the source
  int64_t n = 0;                  ///< vertices of the trees it was last built from
  int64_t cap = 0;                ///< vertices `prev` holds per tree
  dyng_hip::dbuf<int32_t> prev;   ///< those trees' parents, tree k at k * cap
  dyng_hip::dbuf<const int32_t*> old_ptrs, new_ptrs;
  dyng_hip::dbuf<int32_t> dterms;
  dyng_hip::dbuf<int32_t> l0, l1, l2, l3, l4;  ///< the edge list, then the batches
  dyng_hip::dbuf<unsigned long long> totals;
  dyng_hip_mosp() = default;
  dyng_hip_mosp(const dyng_hip_mosp&) = delete;
  dyng_hip_mosp& operator=(const dyng_hip_mosp&) = delete;
  ~dyng_hip_mosp() {
    if (cr != nullptr) (void)dyng_hip_result_destroy(cr);
    if (cg != nullptr) (void)dyng_hip_graph_destroy(cg);
  }
};

namespace {

/// Keep the trees' parents for the next diff.
void mosp_save_trees(dyng_hip_mosp* m, const int32_t* const* tree_parents, int64_t n, hipStream_t s) {
  if (n > m->cap) {
    m->cap = std::max<int64_t>(n, m->cap + m->cap / 2);
    m->prev.reserve(static_cast<size_t>(m->cap) * m->K);
  }
  for (int k = 0; k < m->K; ++k) {
    if (n > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(m->prev.p + static_cast<int64_t>(k) * m->cap, tree_parents[k], n * sizeof(int32_t),
                                  hipMemcpyDeviceToDevice, s));
    }
  }
  m->n = n;
}

void mosp_stats_of(const dyng_hip_mosp* m, dyng_hip_mosp_stats* out) {
  out->preference_scale = m->terms.scale;
  out->base = m->terms.base;
  if (m->cg != nullptr) {
    out->combined_edges = m->cg->m;
    out->weight_sum = m->cg->wsum.empty() ? 0 : m->cg->wsum[0];
  }
}

}  // namespace

extern "C" int dyng_hip_mosp_create(const dyng_hip_graph* g, int32_t source, int32_t num_objectives,
                                    const int32_t* const* tree_parents, const int32_t* preferences, void* stream,
                                    dyng_hip_mosp** out, dyng_hip_mosp_stats* stats) {
  return guarded([&] {
    if (out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: no output handle");
    auto m = std::make_unique<dyng_hip_mosp>();
    m->terms = mosp_check(g, source, num_objectives, tree_parents, preferences);
    m->device = g->device;
    m->source = source;
    m->K = num_objectives;
    const int64_t n = g->n;
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    m->dterms.reserve(num_objectives);
    m->new_ptrs.reserve(num_objectives);
    m->old_ptrs.reserve(num_objectives);
    m->totals.reserve(4);
    DYNG_HIP_TRY(hipStreamSynchronize(s));  // the trees are final before the pointer copies below
    DYNG_HIP_TRY(hipMemcpy(m->dterms.p, m->terms.terms.data(), num_objectives * sizeof(int32_t), hipMemcpyHostToDevice));
    DYNG_HIP_TRY(hipMemcpy(m->new_ptrs.p, tree_parents, num_objectives * sizeof(const int32_t*), hipMemcpyHostToDevice));
    const int64_t list_cap = std::max<int64_t>(1, n * num_objectives);
    m->l0.reserve(list_cap);
    m->l1.reserve(list_cap);
    m->l2.reserve(list_cap);
    int64_t weight_sum = 0;
    const int64_t edges =
        n > 0 ? mosp_combined_list(m->new_ptrs.p, num_objectives, n, source, static_cast<int32_t>(m->terms.base),
                                   m->dterms.p, m->l0.p, m->l1.p, m->l2.p, m->totals.p, s, &weight_sum)
              : 0;
    // a simple graph with sorted rows that keeps MOSP's batch rules (upserts, deletions first) and
    // grows with the trees
    const dyng_hip_graph_props props{1, 1, 0, 0, 0, 0, 1, 1, 0.25};
    call(dyng_hip_graph_create_from_edges(g->device, &props, n, edges, 1, m->l0.p, m->l1.p, m->l2.p, 1, s, &m->cg));
    if (n > 0) {
      const dyng_hip_options opt{0, 0, DYNG_HIP_ENGINE_AUTOMATIC, 1};
      call(dyng_hip_api_compute(m->cg, source, &opt, s, &m->cr));
    }
    mosp_save_trees(m.get(), tree_parents, n, s);
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    if (stats != nullptr) {
      *stats = dyng_hip_mosp_stats{};
      mosp_stats_of(m.get(), stats);
      stats->weight_sum = weight_sum;
    }
    *out = m.release();
  });
}

extern "C" int dyng_hip_mosp_update(dyng_hip_mosp* m, const dyng_hip_graph* g, const int32_t* const* tree_parents,
                                    void* stream, dyng_hip_mosp_stats* stats) {
  return guarded([&] {
    if (m == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: null state");
    (void)mosp_check(g, m->source, m->K, tree_parents, nullptr);
    if (g->device != m->device) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the graph is on another device");
    const int64_t n_old = m->n, n_new = g->n;
    if (n_new < n_old) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the trees have fewer vertices than before");
    scoped_device guard(g->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    dyng_hip_mosp_stats st{};
    if (n_new == 0) {
      mosp_stats_of(m, &st);
      if (stats != nullptr) *stats = st;
      return;
    }
    std::vector<const int32_t*> olds(static_cast<size_t>(m->K));
    for (int k = 0; k < m->K; ++k) olds[k] = m->prev.p + static_cast<int64_t>(k) * m->cap;
    DYNG_HIP_TRY(hipStreamSynchronize(s));  // the repaired trees are final; the pointer arrays are free
    DYNG_HIP_TRY(hipMemcpy(m->new_ptrs.p, tree_parents, m->K * sizeof(const int32_t*), hipMemcpyHostToDevice));
    DYNG_HIP_TRY(hipMemcpy(m->old_ptrs.p, olds.data(), m->K * sizeof(const int32_t*), hipMemcpyHostToDevice));
    const int64_t ops_cap = std::max<int64_t>(1, std::max(n_old, n_new) * m->K);
    for (auto* b : {&m->l0, &m->l1, &m->l2, &m->l3, &m->l4}) b->reserve(ops_cap);
    int64_t nd = 0, ni = 0;
    mosp_combined_diff(m->old_ptrs.p, n_old, m->new_ptrs.p, n_new, m->K, m->source,
                       static_cast<int32_t>(m->terms.base), m->dterms.p, m->l0.p, m->l1.p, m->l2.p, m->l3.p, m->l4.p,
                       m->totals.p, s, &nd, &ni);
    if (m->cr == nullptr) {  // the first trees had no vertex: build as on create
      fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the combined state was created on an empty graph; create it again");
    }
    // vertices the trees gained: rows of their own in the combined graph even before an edge
    // names them (a new vertex may stay unreached)
    if (n_new > m->cg->n) {
      graph_grow(m->cg, n_new, s);
      grow_result(m->cr, n_new, s);
    }
    // the change between the two combined graphs, applied and repaired like any batch
    dyng_hip_batch b{m->l2.p, m->l3.p, m->l4.p, ni, m->l0.p, m->l1.p, nd, 1, 1};
    dyng_hip_stats ss{};
    call(dyng_hip_api_update(m->cg, &b, m->cr, s, &ss));
    mosp_save_trees(m, tree_parents, n_new, s);
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    mosp_stats_of(m, &st);
    st.affected = ss.affected;
    st.deletions = nd;
    st.insertions = ni;
    st.solve = ss.device;
    if (stats != nullptr) *stats = st;
  });
}

extern "C" int dyng_hip_mosp_arrays(const dyng_hip_mosp* m, const int64_t** distances, const int32_t** parents,
                                    int64_t* n) {
  return guarded([&] {
    if (m == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: null state");
    if (m->cr == nullptr) {
      if (distances) *distances = nullptr;
      if (parents) *parents = nullptr;
      if (n) *n = 0;
      return;
    }
    if (distances) *distances = m->cr->distances.p;
    if (parents) *parents = m->cr->parents.p;
    if (n) *n = m->cr->n;
  });
}

extern "C" int dyng_hip_mosp_path_costs(const dyng_hip_mosp* m, const dyng_hip_graph* g, int32_t columns,
                                        int64_t* costs, void* stream) {
  return guarded([&] {
    if (m == nullptr || g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: null state or graph");
    if (columns < 1 || columns > g->K || costs == nullptr) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: columns must be in [1, " + std::to_string(g->K) + "] with an output array");
    }
    if (m->cr == nullptr || g->n == 0) return;
    if (m->cr->n != g->n) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the combined tree does not match the graph");
    scoped_device guard(g->device);
    std::lock_guard<std::mutex> lk(pool().lock);  // the doubling scratch is the device's cache
    mosp_path_costs(g, m->cr->parents.p, g->n, m->source, columns, costs, static_cast<hipStream_t>(stream));
  });
}

extern "C" int dyng_hip_mosp_destroy(dyng_hip_mosp* m) {
  return guarded([&] { delete m; });
}
