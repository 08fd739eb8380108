// The hypergraph container of hyper_sssp (host side, C++): weighted hypergraphs with slack_csr
// storage (cpp/include/dyng/hypergraph/hypergraph.hpp), their batch apply with the reference's
// phases, skip rules and ids (cpp/src/hypergraph/hypergraph.cpp:613-760,
// cpp/src/hypergraph/id_allocator.cpp:119-580, cpp/src/hypergraph/capacity.cpp:105-170,
// cpp/src/hypergraph/escher/store_policy.cpp:71-116), and the line graph a result keeps on the
// device: its edges for a solve, and the edge batch that carries it across one apply.
#include "internal.hpp"

#include <algorithm>
#include <limits>
#include <climits>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace dyng_hip {
namespace {

constexpr int64_t kIntMax = INT32_MAX;

[[noreturn]] void capacity_fail(const std::string& m) { throw status_error(DYNG_HIP_CAPACITY, m); }

int64_t padded4(int64_t c) { return ((c + 4) / 4) * 4; }                                    // rules 1, 3
int64_t aligned(int64_t c, int64_t a) { return c <= 0 ? 0 : (c + a - 1) / a * a; }          // rule 2

/// The payload policy of the hyperedge tree (to_store_policy, capacity.cpp:111-145).
struct policy {
  bool grows = true;
  bool compacts = true;
  int64_t alignment = 4;
  int64_t headroom = 8;
  int64_t max_payload = kIntMax;

  static policy of(const dyng_hip_hyper_props& p) {
    policy q;
    q.grows = p.grow_factor != 0.0;
    q.compacts = p.compact_before_grow != 0;
    q.alignment = p.row_alignment > 0 ? p.row_alignment : 4;
    q.headroom = p.compaction_headroom;
    const int64_t limit = p.max_payload > 0 ? std::min<int64_t>(p.max_payload, kIntMax) : kIntMax;
    q.max_payload = q.grows ? limit : kIntMax;
    return q;
  }
};

/// initial_payload (capacity.cpp:147-170).
int64_t initial_payload(const dyng_hip_hyper_props& p, int64_t size) {
  const int64_t limit = p.max_payload > 0 ? std::min<int64_t>(p.max_payload, kIntMax) : kIntMax;
  const double scaled = static_cast<double>(size) * p.initial_headroom;
  const int64_t slack = scaled < static_cast<double>(limit) ? static_cast<int64_t>(scaled) : limit;
  int64_t cap = size + slack + std::min<int64_t>(p.initial_slack, limit);
  cap = std::min(std::max(cap, p.min_initial_payload), limit);
  if (cap < size) {
    capacity_fail("dyng: hypergraph: a tree of " + std::to_string(size) +
                  " payload entries exceeds the largest payload of its capacity policy");
  }
  return cap;
}

enum : uint8_t { kAbsent = 0, kLive = 1, kDead = 2 };

/// The best-fit id allocator of the hyperedge tree (id_allocator.cpp:119-580): per key (id + 1)
/// its state, slot length, value count and tail segment, the bump pointer and the capacity.
/// Values are kept only for rows whose tail is an overflow segment (a later removal leaves in the
/// tail's occupancy what is left there).
struct allocator {
  policy pol;
  std::vector<uint8_t> state{kAbsent};
  std::vector<int32_t> length{0}, count{0}, tail_capacity{0}, occupancy{0};
  std::map<int64_t, std::vector<int32_t>> tails;
  std::set<int64_t> dead_keys;
  int64_t max_key = 0, live = 0, dead = 0, bump = 0, capacity = 0, compactions = 0;

  void construct(const std::vector<int32_t>& counts) {
    int64_t size = 0;
    for (int32_t c : counts) size += padded4(c);
    const size_t n = counts.size();
    state.assign(n + 1, kLive);
    state[0] = kAbsent;
    length.assign(n + 1, 0);
    count.assign(n + 1, 0);
    tail_capacity.assign(n + 1, 0);
    occupancy.assign(n + 1, 0);
    for (size_t i = 0; i < n; ++i) {
      length[i + 1] = static_cast<int32_t>(padded4(counts[i]));
      count[i + 1] = counts[i];
      tail_capacity[i + 1] = length[i + 1] - 1;
      occupancy[i + 1] = counts[i];
    }
    max_key = static_cast<int64_t>(n);
    live = max_key;
    dead = 0;
    bump = size;
  }

  void at(int64_t key) {
    if (static_cast<int64_t>(state.size()) <= key) {
      const size_t s = static_cast<size_t>(key + 1);
      state.resize(s, kAbsent);
      length.resize(s, 0);
      count.resize(s, 0);
      tail_capacity.resize(s, 0);
      occupancy.resize(s, 0);
    }
  }

  void clear(int64_t key) {
    state[key] = kAbsent;
    length[key] = count[key] = tail_capacity[key] = occupancy[key] = 0;
    tails.erase(key);
  }

  int64_t grown(int64_t needed) const {  // store_policy::grown_capacity (store_policy.cpp:100-116)
    if (!pol.grows) {
      capacity_fail("dyng: CBST h2v: " + std::to_string(needed) + " payload entries do not fit its capacity of " +
                    std::to_string(capacity) + " (the policy does not grow the payload)");
    }
    const int64_t g = std::min(std::max(needed + needed / 4, 3 * capacity / 2), pol.max_payload);
    if (needed > g) {
      capacity_fail("dyng: CBST h2v: " + std::to_string(needed) + " payload entries exceed the largest payload");
    }
    return g;
  }

  void ensure(int64_t needed) {
    if (needed > capacity) capacity = grown(needed);
  }

  void keep_headroom(int64_t compactions_before) {  // id_allocator.cpp:212-225
    if (pol.headroom <= 0 || compactions == compactions_before || !pol.grows) return;
    const int64_t slack = capacity / pol.headroom;
    if (capacity - bump >= slack) return;
    ensure(std::min(bump + slack, std::min<int64_t>(pol.max_payload, kIntMax)));
  }

  void compact() {  // id_allocator.cpp:227-263
    if (live + dead == 0) return;
    int64_t total = 0;
    for (int64_t k = 1; k <= max_key; ++k) {
      if (state[k] == kLive) total += padded4(count[k]);
      else if (state[k] == kDead) total += length[k];
    }
    if (total > kIntMax) capacity_fail("dyng: CBST h2v: the compacted rows exceed the int32 offsets");
    ensure(total);
    for (int64_t k = 1; k <= max_key; ++k) {
      if (state[k] == kLive) {
        length[k] = static_cast<int32_t>(padded4(count[k]));
        tail_capacity[k] = length[k] - 1;
        occupancy[k] = count[k];
      } else if (state[k] == kDead) {
        tail_capacity[k] = length[k] - 1;
        occupancy[k] = 0;
      }
    }
    bump = total;
    ++compactions;
    tails.clear();
  }

  void erase(const std::vector<int64_t>& keys) {  // id_allocator.cpp:265-282
    for (int64_t k : keys) {
      if (k > 0 && k < static_cast<int64_t>(state.size()) && state[k] == kLive) {
        state[k] = kDead;
        dead_keys.insert(k);
        --live;
        ++dead;
        tails.erase(k);
      }
    }
  }

  std::vector<int64_t> insert(const std::vector<int32_t>& counts) {  // id_allocator.cpp:284-411
    const size_t num = counts.size();
    std::vector<int64_t> keys(num, 0);
    if (num == 0) return keys;
    const int64_t compactions_before = compactions;
    std::vector<char> matched(num, 0);
    const size_t reuse_k = std::min<size_t>(num, static_cast<size_t>(dead));
    if (reuse_k > 0) {
      std::vector<int64_t> dk(dead_keys.begin(), dead_keys.end());  // key order
      const size_t d = dk.size();
      std::vector<int64_t> cap(d);
      for (size_t j = 0; j < d; ++j) cap[j] = length[dk[j]] - 1;
      std::vector<size_t> slot_order(d), item_order(reuse_k);
      for (size_t j = 0; j < d; ++j) slot_order[j] = j;
      for (size_t i = 0; i < reuse_k; ++i) item_order[i] = i;
      std::stable_sort(slot_order.begin(), slot_order.end(), [&](size_t a, size_t b) { return cap[a] < cap[b]; });
      std::stable_sort(item_order.begin(), item_order.end(),
                       [&](size_t a, size_t b) { return counts[a] < counts[b]; });
      std::vector<int64_t> sorted_cap(d);
      for (size_t j = 0; j < d; ++j) sorted_cap[j] = cap[slot_order[j]];
      int64_t prefix_max = LLONG_MIN;
      for (size_t i = 0; i < reuse_k; ++i) {
        const size_t item = item_order[i];
        const int64_t lo = std::lower_bound(sorted_cap.begin(), sorted_cap.end(), int64_t{counts[item]}) -
                           sorted_cap.begin();
        prefix_max = std::max(prefix_max, lo - static_cast<int64_t>(i));
        const int64_t assigned = static_cast<int64_t>(i) + prefix_max;
        if (assigned >= static_cast<int64_t>(d)) break;
        const int64_t key = dk[slot_order[static_cast<size_t>(assigned)]];
        state[key] = kLive;
        dead_keys.erase(key);
        count[key] = counts[item];
        tail_capacity[key] = length[key] - 1;
        occupancy[key] = counts[item];
        --dead;
        ++live;
        keys[item] = key;
        matched[item] = 1;
      }
    }
    std::vector<size_t> surplus;
    for (size_t i = 0; i < num; ++i) {
      if (!matched[i]) surplus.push_back(i);
    }
    if (!surplus.empty()) {
      int64_t slots = 0;
      for (size_t i : surplus) slots += aligned(counts[i], pol.alignment) + 1;
      if (bump + slots > capacity && pol.compacts) compact();
      ensure(bump + slots);
      bump += slots;
      // the rebuild: deleted slots not reused are dropped; the surplus rows take the keys after
      // the last live key, in item order
      for (int64_t k : dead_keys) clear(k);
      dead_keys.clear();
      dead = 0;
      int64_t last_live = max_key;
      while (last_live > 0 && state[last_live] != kLive) --last_live;
      int64_t next = last_live + 1;
      for (size_t i : surplus) {
        const int64_t a = aligned(counts[i], pol.alignment);
        at(next);
        tails.erase(next);
        state[next] = kLive;
        length[next] = static_cast<int32_t>(a + 1);
        count[next] = counts[i];
        tail_capacity[next] = static_cast<int32_t>(a);
        occupancy[next] = counts[i];
        ++live;
        keys[i] = next;
        ++next;
      }
      for (int64_t k = next; k <= max_key; ++k) clear(k);
      max_key = next - 1;
    }
    keep_headroom(compactions_before);
    return keys;
  }

  using groups = std::vector<std::pair<int64_t, std::vector<int32_t>>>;

  void fill(const groups& g, bool allow_compact) {  // id_allocator.cpp:413-490
    if (g.empty()) return;
    std::vector<int64_t> taken(g.size(), 0), segment(g.size(), 0);
    int64_t total = 0;
    for (size_t i = 0; i < g.size(); ++i) {
      const int64_t key = g[i].first;
      const auto& values = g[i].second;
      const int64_t occ = std::min<int64_t>(std::max<int32_t>(occupancy[key], 0), tail_capacity[key]);
      const int64_t room = tail_capacity[key] - occ;
      const int64_t k = static_cast<int64_t>(values.size());
      const int64_t fit = std::min(k, room);
      auto tail = tails.find(key);
      if (tail != tails.end()) tail->second.insert(tail->second.end(), values.begin(), values.begin() + fit);
      if (k <= room) {
        occupancy[key] = static_cast<int32_t>(occ + k);
        count[key] += static_cast<int32_t>(k);
        taken[i] = k;
      } else {
        occupancy[key] = tail_capacity[key];
        count[key] += static_cast<int32_t>(room);
        taken[i] = room;
        segment[i] = padded4(k - room);
        total += segment[i];
      }
    }
    const int64_t needed = bump + total;
    if (total > 0 && needed > capacity && allow_compact && pol.compacts) {
      groups rest;
      for (size_t i = 0; i < g.size(); ++i) {
        if (segment[i] != 0) rest.emplace_back(g[i].first, std::vector<int32_t>(g[i].second.begin() + taken[i],
                                                                                   g[i].second.end()));
      }
      compact();
      fill(rest, false);
      return;
    }
    if (total > 0) {
      ensure(needed);
      for (size_t i = 0; i < g.size(); ++i) {
        if (segment[i] == 0) continue;
        const int64_t key = g[i].first;
        const int64_t k = static_cast<int64_t>(g[i].second.size());
        count[key] += static_cast<int32_t>(k - taken[i]);
        tail_capacity[key] = static_cast<int32_t>(segment[i] - 1);
        occupancy[key] = static_cast<int32_t>(k - taken[i]);
        tails[key].assign(g[i].second.begin() + taken[i], g[i].second.end());
      }
      bump = needed;
    }
  }

  void append(const groups& g) {  // id_allocator.cpp:492-498
    const int64_t before = compactions;
    fill(g, true);
    keep_headroom(before);
  }

  void remove_values(const groups& g) {  // id_allocator.cpp:500-540
    for (const auto& [key, values] : g) {
      if (key <= 0 || key >= static_cast<int64_t>(state.size()) || state[key] != kLive) continue;
      std::vector<int32_t> removed(values);
      std::sort(removed.begin(), removed.end());
      removed.erase(std::unique(removed.begin(), removed.end()), removed.end());
      count[key] -= static_cast<int32_t>(removed.size());
      auto tail = tails.find(key);
      if (tail == tails.end()) {
        occupancy[key] = count[key];
      } else {
        auto& tv = tail->second;
        tv.erase(std::remove_if(tv.begin(), tv.end(),
                                [&](int32_t v) { return std::binary_search(removed.begin(), removed.end(), v); }),
                 tv.end());
        occupancy[key] = static_cast<int32_t>(tv.size());
      }
    }
  }
};

void remove_one(std::vector<int64_t>& list, int64_t h) {  // the first occurrence, swap-removed
  for (size_t i = 0; i < list.size(); ++i) {
    if (list[i] == h) {
      list[i] = list.back();
      list.pop_back();
      return;
    }
  }
}

}  // namespace
}  // namespace dyng_hip

/// The container behind a hypergraph handle.
struct dyng_hip_hypergraph {
  int32_t n = 0;
  bool weighted = false;
  int32_t reuse = 0;  ///< 0 erase_first, 1 insert_first
  dyng_hip_hyper_props props{};
  std::vector<std::vector<int32_t>> rows;  ///< members by id, sorted; empty: dead
  std::vector<int64_t> w;                  ///< weight by id (0 for a dead id)
  std::vector<std::vector<int64_t>> v2h;   ///< the live hyperedges of each vertex (unsorted)
  int64_t live = 0;
  int64_t version = 0;
  dyng_hip::allocator alloc;
  // the last apply: the ids whose rows changed, their rows and weights before it
  std::vector<int64_t> touched;
  std::vector<std::vector<int32_t>> pre_rows;
  std::vector<int64_t> pre_w;
  // scratch and the arrays the last query returned (valid until the next call on the handle)
  /// Marks of one id, side by side so that a visited id costs one cache line: `a` and `b` hold
  /// the stamps of the two neighbour sets being built, `t` the touched-id stamp.
  struct id_marks {
    int64_t a = -1, b = -1, t = -1;
  };
  std::vector<id_marks> marks;
  int64_t stamp = 0;  ///< marks below it are stale: each query takes fresh values above it
  std::vector<std::pair<int32_t, int64_t>> pre_pairs;  ///< (vertex, touched id) before the apply

  /// The marks at least b ids long; entries past their old length start stale.
  void marks_for(int64_t b) {
    const size_t need = static_cast<size_t>(std::max<int64_t>(b, 1));
    if (marks.size() < need) marks.resize(need);
  }
  std::vector<int32_t> out_src, out_dst, out_w, out_del_src, out_del_dst;
  std::vector<int64_t> out_offsets;
  std::vector<int32_t> out_members;

  int64_t id_bound() const { return alloc.max_key; }

  void set_row(int64_t h, std::vector<int32_t> row) {
    if (static_cast<int64_t>(rows.size()) <= h) {
      rows.resize(static_cast<size_t>(h + 1));
      w.resize(static_cast<size_t>(h + 1), 0);
    }
    auto& old = rows[static_cast<size_t>(h)];
    live += (row.empty() ? 0 : 1) - (old.empty() ? 0 : 1);
    for (int32_t v : old) dyng_hip::remove_one(v2h[static_cast<size_t>(v)], h);
    for (int32_t v : row) v2h[static_cast<size_t>(v)].push_back(h);
    if (row.empty()) w[static_cast<size_t>(h)] = 0;  // a dead id has no weight
    old = std::move(row);
  }

  /// The live ids sharing a vertex with h (rows as they are now), marked with `st` in `field`.
  template <class Visit>
  void neighbours(const std::vector<int32_t>& row, int64_t h, int64_t id_marks::*field, int64_t st, Visit visit) {
    for (int32_t v : row) {
      for (int64_t g : v2h[static_cast<size_t>(v)]) {
        int64_t& m = marks[static_cast<size_t>(g)].*field;
        if (g != h && m != st) {
          m = st;
          visit(g);
        }
      }
    }
  }
};

using dyng_hip::fail;
using dyng_hip::guarded;

namespace {

std::vector<int32_t> sorted_unique(const int32_t* a, int64_t n) {
  std::vector<int32_t> r(a, a + n);
  std::sort(r.begin(), r.end());
  r.erase(std::unique(r.begin(), r.end()), r.end());
  return r;
}

}  // namespace

extern "C" int dyng_hip_hypergraph_create(const int64_t* offsets, const int32_t* members, int64_t num_hyperedges,
                                          const int64_t* weights, int32_t num_vertices,
                                          const dyng_hip_hyper_props* props, dyng_hip_hypergraph** out) {
  return guarded([&] {
    if (out == nullptr || props == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: null argument");
    if (num_hyperedges < 0 || num_vertices < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: negative size");
    if (num_hyperedges > 0 && offsets == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: null offsets");
    auto h = std::make_unique<dyng_hip_hypergraph>();
    h->n = num_vertices;
    h->weighted = weights != nullptr;
    h->reuse = props->id_reuse;
    h->props = *props;
    h->alloc.pol = dyng_hip::policy::of(*props);
    h->v2h.resize(static_cast<size_t>(num_vertices));
    std::vector<int32_t> counts(static_cast<size_t>(num_hyperedges));
    h->rows.resize(static_cast<size_t>(num_hyperedges));
    h->w.assign(static_cast<size_t>(num_hyperedges), 0);
    for (int64_t i = 0; i < num_hyperedges; ++i) {
      if (offsets[i + 1] < offsets[i]) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: the offsets decrease");
      auto row = sorted_unique(members + offsets[i], offsets[i + 1] - offsets[i]);
      if (!row.empty() && (row.front() < 0 || row.back() >= num_vertices)) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: a member is outside [0, " + std::to_string(num_vertices) + ")");
      }
      if (weights != nullptr && weights[i] < 1) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: weights are >= 1");
      counts[static_cast<size_t>(i)] = static_cast<int32_t>(row.size());
      if (weights != nullptr) h->w[static_cast<size_t>(i)] = weights[i];
      for (int32_t v : row) h->v2h[static_cast<size_t>(v)].push_back(i);
      h->live += row.empty() ? 0 : 1;
      h->rows[static_cast<size_t>(i)] = std::move(row);
    }
    int64_t size = 0;
    for (int32_t c : counts) size += dyng_hip::padded4(c);
    h->alloc.capacity = dyng_hip::initial_payload(*props, size);
    h->alloc.construct(counts);
    *out = h.release();
  });
}

extern "C" int dyng_hip_hypergraph_destroy(dyng_hip_hypergraph* h) {
  delete h;
  return DYNG_HIP_OK;
}

extern "C" int dyng_hip_hypergraph_clone(const dyng_hip_hypergraph* h, dyng_hip_hypergraph** out) {
  return guarded([&] {
    if (h == nullptr || out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: null argument");
    *out = new dyng_hip_hypergraph(*h);
  });
}

extern "C" int dyng_hip_hypergraph_info(const dyng_hip_hypergraph* h, int32_t* num_vertices, int64_t* num_live,
                                        int64_t* id_bound, int64_t* version, int32_t* weighted) {
  return guarded([&] {
    if (h == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: null handle");
    if (num_vertices) *num_vertices = h->n;
    if (num_live) *num_live = h->live;
    if (id_bound) *id_bound = h->id_bound();
    if (version) *version = h->version;
    if (weighted) *weighted = h->weighted ? 1 : 0;
  });
}

extern "C" int dyng_hip_hypergraph_rows(dyng_hip_hypergraph* h, const int64_t** offsets, const int32_t** members,
                                        const int64_t** weights) {
  return guarded([&] {
    if (h == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph: null handle");
    const int64_t b = h->id_bound();
    h->out_offsets.assign(static_cast<size_t>(b + 1), 0);
    h->out_members.clear();
    for (int64_t i = 0; i < b; ++i) {
      if (i < static_cast<int64_t>(h->rows.size())) {
        const auto& r = h->rows[static_cast<size_t>(i)];
        h->out_members.insert(h->out_members.end(), r.begin(), r.end());
      }
      h->out_offsets[static_cast<size_t>(i + 1)] = static_cast<int64_t>(h->out_members.size());
    }
    if (static_cast<int64_t>(h->w.size()) < b) h->w.resize(static_cast<size_t>(b), 0);
    if (offsets) *offsets = h->out_offsets.data();
    if (members) *members = h->out_members.data();
    if (weights) *weights = h->weighted ? h->w.data() : nullptr;
  });
}

extern "C" int dyng_hip_hypergraph_apply(dyng_hip_hypergraph* h, const dyng_hip_hyper_batch* b, int64_t max_weight,
                                         dyng_hip_hyper_summary* out, int64_t* inserted_ids) {
  return guarded([&] {
    if (h == nullptr || b == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph::apply: null argument");
    // ---- checks before anything changes ----
    std::vector<std::vector<int32_t>> rows_in(static_cast<size_t>(b->num_insertions));
    int64_t dropped = 0;
    for (int64_t i = 0; i < b->num_insertions; ++i) {
      const int64_t len = b->insert_offsets[i + 1] - b->insert_offsets[i];
      if (len <= 0) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph::apply: an inserted hyperedge has no member");
      auto row = sorted_unique(b->insert_members + b->insert_offsets[i], len);
      if (row.front() < 0 || row.back() >= h->n) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph::apply: a member of an inserted hyperedge is outside [0, " +
                                            std::to_string(h->n) + ")");
      }
      dropped += len - static_cast<int64_t>(row.size());
      rows_in[static_cast<size_t>(i)] = std::move(row);
    }
    if (h->weighted && b->num_insertions > 0) {
      if (b->insert_weights == nullptr) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph::apply: a weighted hypergraph needs one insert weight per "
                                        "inserted hyperedge");
      }
      for (int64_t i = 0; i < b->num_insertions; ++i) {
        if (b->insert_weights[i] < 1) fail(DYNG_HIP_INVALID_ARGUMENT, "hypergraph::apply: insert weights are >= 1");
        if (max_weight > 0 && b->insert_weights[i] > max_weight) {
          fail(DYNG_HIP_INVALID_ARGUMENT, "hyper_sssp::update: an inserted weight is outside [1, 2^28]");
        }
      }
    }
    const int64_t bound = h->id_bound();
    auto live = [&](int64_t id) {
      return id >= 0 && id < bound && id < static_cast<int64_t>(h->rows.size()) &&
             !h->rows[static_cast<size_t>(id)].empty();
    };
    // deletions against the hypergraph before the batch (hypergraph.cpp:855-889)
    std::vector<int64_t> deleted;
    std::set<int64_t> seen;
    int64_t ignored = 0;
    for (int64_t i = 0; i < b->num_deletions; ++i) {
      const int64_t id = b->delete_hyperedges[i];
      if (live(id) && seen.insert(id).second) deleted.push_back(id);
      else ++ignored;
    }
    // incidence edits (hypergraph.cpp:613-760): per id in increasing order, its removals then its
    // additions in batch order, each against the row the earlier edits left
    const int64_t nrem = b->num_removals, nadd = b->num_additions;
    std::vector<std::pair<int64_t, int64_t>> words;  // (id, edit)
    for (int64_t j = 0; j < nrem + nadd; ++j) {
      const int64_t id = j < nrem ? b->delete_incidence_hyperedge[j] : b->insert_incidence_hyperedge[j - nrem];
      const int64_t v = j < nrem ? b->delete_incidence_vertex[j] : b->insert_incidence_vertex[j - nrem];
      if (id >= 0 && id < bound && v >= 0 && v < h->n) words.emplace_back(id, j);
    }
    std::stable_sort(words.begin(), words.end(), [](const auto& a, const auto& c) { return a.first < c.first; });
    dyng_hip::allocator::groups removals, additions;
    std::vector<std::pair<int64_t, std::vector<int32_t>>> edited;
    int64_t removed = 0, added = 0;
    for (size_t s = 0; s < words.size();) {
      size_t e = s;
      while (e < words.size() && words[e].first == words[s].first) ++e;
      const int64_t id = words[s].first;
      if (live(id) && seen.count(id) == 0) {
        std::vector<int32_t> row = h->rows[static_cast<size_t>(id)];
        std::vector<int32_t> rem, add;
        for (size_t x = s; x < e; ++x) {
          const int64_t j = words[x].second;
          const int32_t v = static_cast<int32_t>(j < nrem ? b->delete_incidence_vertex[j]
                                                          : b->insert_incidence_vertex[j - nrem]);
          auto it = std::lower_bound(row.begin(), row.end(), v);
          const bool member = it != row.end() && *it == v;
          if (j < nrem) {
            if (!member || row.size() == 1) continue;
            row.erase(it);
            rem.push_back(v);
            ++removed;
          } else {
            if (member) continue;
            row.insert(it, v);
            add.push_back(v);
            ++added;
          }
        }
        if (!rem.empty() || !add.empty()) edited.emplace_back(id, std::move(row));
        if (!rem.empty()) removals.emplace_back(id + 1, std::move(rem));
        if (!add.empty()) additions.emplace_back(id + 1, std::move(add));
      }
      s = e;
    }
    const int64_t skipped_edits = nrem + nadd - removed - added;
    // ---- the store phase in the order of id_reuse; the allocator commits only if every step does.
    // Its steps fail only when a payload would pass the policy's limits (a payload it may not
    // grow, or one past max_payload or the int32 offsets). No step of one apply asks for more than
    // the payload before it, again (a row relocated whole), plus four slots per value of the batch
    // and a row's padding: when a growing policy has room for twice the payload, the capacity and
    // that, the steps run
    // in place; otherwise on a copy committed at the end, so a failed apply changes nothing.
    std::vector<int32_t> counts(rows_in.size());
    int64_t batch_values = 0;
    for (size_t i = 0; i < rows_in.size(); ++i) {
      counts[i] = static_cast<int32_t>(rows_in[i].size());
      batch_values += counts[i];
    }
    for (const auto& a : additions) batch_values += static_cast<int64_t>(a.second.size());
    const dyng_hip::allocator& now = h->alloc;
    const int64_t room = std::min<int64_t>(now.pol.max_payload, dyng_hip::kIntMax);
    const bool in_place =
        now.pol.grows && 2 * now.bump + now.capacity + 4 * batch_values +
                             now.pol.alignment * static_cast<int64_t>(counts.size() + additions.size()) + 65536 <
                         room;
    dyng_hip::allocator copy;
    if (!in_place) copy = h->alloc;
    dyng_hip::allocator& alloc = in_place ? h->alloc : copy;
    std::vector<int64_t> del_keys;
    for (int64_t id : deleted) del_keys.push_back(id + 1);
    std::vector<int64_t> keys;
    if (h->reuse == 1) {
      keys = alloc.insert(counts);
      alloc.erase(del_keys);
      alloc.remove_values(removals);
      alloc.append(additions);
    } else {
      alloc.erase(del_keys);
      alloc.remove_values(removals);
      alloc.append(additions);
      keys = alloc.insert(counts);
    }
    // ---- commit: the rows, the weights, the delta of this apply ----
    if (!in_place) h->alloc = std::move(copy);
    h->touched.clear();
    h->pre_rows.clear();
    h->pre_w.clear();
    // the ids whose rows change, once each in increasing order, with their rows and weights
    // before the apply
    for (int64_t id : deleted) h->touched.push_back(id);
    for (const auto& e : edited) h->touched.push_back(e.first);
    for (int64_t k : keys) h->touched.push_back(k - 1);
    std::sort(h->touched.begin(), h->touched.end());
    h->touched.erase(std::unique(h->touched.begin(), h->touched.end()), h->touched.end());
    for (int64_t id : h->touched) {
      const bool has = id < static_cast<int64_t>(h->rows.size());
      h->pre_rows.push_back(has ? h->rows[static_cast<size_t>(id)] : std::vector<int32_t>{});
      h->pre_w.push_back(has ? h->w[static_cast<size_t>(id)] : 0);
    }
    for (int64_t id : deleted) h->set_row(id, {});
    for (auto& e : edited) h->set_row(e.first, std::move(e.second));
    for (size_t i = 0; i < keys.size(); ++i) {
      const int64_t id = keys[i] - 1;
      h->set_row(id, std::move(rows_in[i]));
      h->w[static_cast<size_t>(id)] = h->weighted ? b->insert_weights[i] : 0;
      if (inserted_ids) inserted_ids[i] = id;
    }
    ++h->version;
    if (out) {
      *out = dyng_hip_hyper_summary{};
      out->inserted_hyperedges = b->num_insertions;
      out->deleted_hyperedges = static_cast<int64_t>(deleted.size());
      out->inserted_incidences = added;
      out->deleted_incidences = removed;
      out->ignored_deletions = ignored;
      out->dropped_members = dropped;
      out->skipped_ops = ignored + dropped + skipped_edits;
      out->num_hyperedges_after = h->live;
      out->id_bound_after = h->id_bound();
    }
  });
}

extern "C" int dyng_hip_hyper_line_graph(dyng_hip_hypergraph* h, int32_t source, int64_t* count, const int32_t** src,
                                         const int32_t** dst, const int32_t** weight) {
  return guarded([&] {
    if (h == nullptr || count == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hyper_sssp: null argument");
    h->out_src.clear();
    h->out_dst.clear();
    h->out_w.clear();
    const int64_t b = h->id_bound();
    h->marks_for(b);
    const int64_t base = h->stamp;  // id t marks its neighbours with base + t
    h->stamp += b + 1;
    if (source >= 0 && source < h->n) {
      std::vector<int64_t> seeds = h->v2h[static_cast<size_t>(source)];
      std::sort(seeds.begin(), seeds.end());
      for (int64_t t : seeds) {
        h->out_src.push_back(0);
        h->out_dst.push_back(static_cast<int32_t>(t + 1));
        h->out_w.push_back(static_cast<int32_t>(h->w[static_cast<size_t>(t)]));
      }
    }
    for (int64_t t = 0; t < b && t < static_cast<int64_t>(h->rows.size()); ++t) {
      const auto& row = h->rows[static_cast<size_t>(t)];
      if (row.empty()) continue;
      const int32_t wt = static_cast<int32_t>(h->w[static_cast<size_t>(t)]);
      h->neighbours(row, t, &dyng_hip_hypergraph::id_marks::a, base + t, [&](int64_t g) {  // the in-edges (g, t): weight w(t)
        h->out_src.push_back(static_cast<int32_t>(g + 1));
        h->out_dst.push_back(static_cast<int32_t>(t + 1));
        h->out_w.push_back(wt);
      });
    }
    *count = static_cast<int64_t>(h->out_src.size());
    if (src) *src = h->out_src.data();
    if (dst) *dst = h->out_dst.data();
    if (weight) *weight = h->out_w.data();
  });
}

extern "C" int dyng_hip_hyper_line_graph_delta(dyng_hip_hypergraph* h, int32_t source, dyng_hip_hyper_delta* out) {
  return guarded([&] {
    if (h == nullptr || out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "hyper_sssp: null argument");
    h->out_src.clear();
    h->out_dst.clear();
    h->out_w.clear();
    h->out_del_src.clear();
    h->out_del_dst.clear();
    const int64_t b = std::max<int64_t>(h->id_bound(), static_cast<int64_t>(h->rows.size()));
    // marks of this query take values above every earlier one (nothing is cleared): id i of the
    // touched list marks with base + 2i and base + 2i + 1, the touched ids with base + 2k
    h->marks_for(b);
    const int64_t base = h->stamp;
    const int64_t touched_mark = base + 2 * static_cast<int64_t>(h->touched.size());
    h->stamp = touched_mark + 1;
    auto& marks = h->marks;
    auto is_touched = [&](int64_t g) { return marks[static_cast<size_t>(g)].t == touched_mark; };
    for (int64_t t : h->touched) marks[static_cast<size_t>(t)].t = touched_mark;
    // the touched ids that held each vertex before the apply: (vertex, id) pairs by vertex
    h->pre_pairs.clear();
    for (size_t i = 0; i < h->touched.size(); ++i) {
      for (int32_t v : h->pre_rows[i]) h->pre_pairs.emplace_back(v, h->touched[i]);
    }
    std::sort(h->pre_pairs.begin(), h->pre_pairs.end());
    auto ins = [&](int64_t a, int64_t c, int64_t wt) {
      h->out_src.push_back(static_cast<int32_t>(a));
      h->out_dst.push_back(static_cast<int32_t>(c));
      h->out_w.push_back(static_cast<int32_t>(wt));
    };
    auto del = [&](int64_t a, int64_t c) {
      h->out_del_src.push_back(static_cast<int32_t>(a));
      h->out_del_dst.push_back(static_cast<int32_t>(c));
    };
    int64_t gained = 0, lost = 0;
    std::vector<int64_t> before, after;
    for (size_t i = 0; i < h->touched.size(); ++i) {
      const int64_t t = h->touched[i];
      const auto& pre = h->pre_rows[i];
      const auto& post = h->rows[static_cast<size_t>(t)];
      const int64_t w_pre = h->pre_w[i];
      const int64_t w_post = h->w[static_cast<size_t>(t)];
      const int64_t sa = base + 2 * static_cast<int64_t>(i), sb = sa + 1;
      // before: the untouched holders of its old members (unchanged by the apply) and the touched
      // ids that held them before; after: today's holders of its members
      before.clear();
      after.clear();
      for (int32_t v : pre) {
        for (int64_t g : h->v2h[static_cast<size_t>(v)]) {
          auto& m = marks[static_cast<size_t>(g)];
          if (g != t && m.t != touched_mark && m.a != sa) {
            m.a = sa;
            before.push_back(g);
          }
        }
        auto it = std::lower_bound(h->pre_pairs.begin(), h->pre_pairs.end(),
                                   std::make_pair(v, std::numeric_limits<int64_t>::min()));
        for (; it != h->pre_pairs.end() && it->first == v; ++it) {
          const int64_t g = it->second;
          auto& m = marks[static_cast<size_t>(g)];
          if (g != t && m.a != sa) {
            m.a = sa;
            before.push_back(g);
          }
        }
      }
      h->neighbours(post, t, &dyng_hip_hypergraph::id_marks::b, sb, [&](int64_t g) { after.push_back(g); });
      for (int64_t g : before) {
        if (marks[static_cast<size_t>(g)].b == sb) continue;  // kept
        if (!is_touched(g) || t < g) ++lost;
        del(g + 1, t + 1);
        if (!is_touched(g)) del(t + 1, g + 1);  // a touched g deletes it as its in-edge
      }
      for (int64_t g : after) {
        if (marks[static_cast<size_t>(g)].a == sa) {
          if (w_pre != w_post) ins(g + 1, t + 1, w_post);  // kept, the head's weight changed
          continue;
        }
        if (!is_touched(g) || t < g) ++gained;
        ins(g + 1, t + 1, w_post);
        if (!is_touched(g)) ins(t + 1, g + 1, h->w[static_cast<size_t>(g)]);
      }
      const bool was = std::binary_search(pre.begin(), pre.end(), source);
      const bool now = std::binary_search(post.begin(), post.end(), source);
      if (was && !now) del(0, t + 1);
      else if (now && (!was || w_pre != w_post)) ins(0, t + 1, w_post);
    }
    out->num_insertions = static_cast<int64_t>(h->out_src.size());
    out->insert_src = h->out_src.data();
    out->insert_dst = h->out_dst.data();
    out->insert_weights = h->out_w.data();
    out->num_deletions = static_cast<int64_t>(h->out_del_src.size());
    out->delete_src = h->out_del_src.data();
    out->delete_dst = h->out_del_dst.data();
    out->line_graph_insertions = gained;
    out->line_graph_deletions = lost;
  });
}
