// Synthetic inputs of the benchmark configurations (host, OpenMP): Erdos-Renyi, RMAT and 2-D grid
// graphs, and mixed batches. Every random draw is splitmix64 of (seed, draw index), so a draw does
// not depend on the thread count or on how many draws came before it: the CPU and GPU legs and
// every rerun see the same graph and batches.
//
// "m unique edges" means: draw candidates (u, v) in draw order, drop self-loops, keep the first m
// distinct pairs by draw index; weights are drawn per kept edge from its draw index.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <parallel/algorithm>
#include <unordered_set>
#include <vector>

#include <omp.h>

namespace {

inline uint64_t splitmix(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

inline double unit(uint64_t r) { return static_cast<double>(r >> 11) * (1.0 / 9007199254740992.0); }

constexpr uint64_t kWeightStream = 0x5745494748545321ULL;

struct draw {
  uint64_t key;
  uint64_t index;
};

/// The first m distinct keys by draw index, sorted by key; false if fewer than m are distinct.
bool first_unique(std::vector<draw>& d, int64_t m, std::vector<draw>& out) {
  __gnu_parallel::sort(d.begin(), d.end(), [](const draw& a, const draw& b) {
    return a.key < b.key || (a.key == b.key && a.index < b.index);
  });
  std::vector<draw> u;
  u.reserve(d.size());
  for (size_t i = 0; i < d.size(); ++i) {
    if (i == 0 || d[i].key != d[i - 1].key) u.push_back(d[i]);
  }
  if (static_cast<int64_t>(u.size()) < m) return false;
  __gnu_parallel::nth_element(u.begin(), u.begin() + m, u.end(),
                              [](const draw& a, const draw& b) { return a.index < b.index; });
  u.resize(static_cast<size_t>(m));
  __gnu_parallel::sort(u.begin(), u.end(), [](const draw& a, const draw& b) { return a.key < b.key; });
  out.swap(u);
  return true;
}

template <typename pair_fn>
int64_t unique_edges(int64_t m, int32_t wmax, uint64_t seed, int32_t* src, int32_t* dst, int32_t* w, pair_fn pair) {
  double factor = 1.05;
  std::vector<draw> kept;
  for (int attempt = 0; attempt < 12; ++attempt) {
    const int64_t D = static_cast<int64_t>(static_cast<double>(m) * factor) + 1024;
    std::vector<draw> d(static_cast<size_t>(D));
    int64_t valid = 0;
#pragma omp parallel for schedule(static) reduction(+ : valid)
    for (int64_t i = 0; i < D; ++i) {
      uint32_t u = 0, v = 0;
      pair(static_cast<uint64_t>(i), u, v);
      // a self-loop becomes a key that sorts last and is never kept (index beyond every draw)
      d[static_cast<size_t>(i)] = u == v ? draw{~0ULL, ~0ULL} : draw{(static_cast<uint64_t>(u) << 32) | v, static_cast<uint64_t>(i)};
      valid += u == v ? 0 : 1;
    }
    if (valid >= m && first_unique(d, m, kept) && kept.back().key != ~0ULL) break;
    kept.clear();
    factor *= 1.4;
  }
  if (static_cast<int64_t>(kept.size()) < m) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; ++e) {
    src[e] = static_cast<int32_t>(kept[e].key >> 32);
    dst[e] = static_cast<int32_t>(kept[e].key & 0xffffffffULL);
    w[e] = 1 + static_cast<int32_t>(splitmix(seed ^ kWeightStream, kept[e].index) % static_cast<uint64_t>(wmax));
  }
  return m;
}

inline void rmat_pair(uint64_t seed, uint64_t i, int scale, double a, double b, double c, uint32_t& u, uint32_t& v) {
  u = 0;
  v = 0;
  for (int l = 0; l < scale; ++l) {
    const double r = unit(splitmix(seed, i * static_cast<uint64_t>(scale) + l));
    const uint32_t bu = r >= a + b ? 1u : 0u;
    const uint32_t bv = (r >= a && r < a + b) || r >= a + b + c ? 1u : 0u;
    u = (u << 1) | bu;
    v = (v << 1) | bv;
  }
}

}  // namespace

extern "C" {

uint64_t dyng_gen_splitmix(uint64_t seed, uint64_t i) { return splitmix(seed, i); }

int64_t dyng_gen_er(int64_t n, int64_t m, int32_t wmax, uint64_t seed, int32_t* src, int32_t* dst, int32_t* w) {
  if (n < 2 || m < 0 || m > n * (n - 1) || wmax < 1) return -1;
  return unique_edges(m, wmax, seed, src, dst, w, [&](uint64_t i, uint32_t& u, uint32_t& v) {
    u = static_cast<uint32_t>(splitmix(seed, 2 * i) % static_cast<uint64_t>(n));
    v = static_cast<uint32_t>(splitmix(seed, 2 * i + 1) % static_cast<uint64_t>(n));
  });
}

int64_t dyng_gen_rmat(int32_t scale, int64_t m, double a, double b, double c, int32_t wmax, uint64_t seed,
                      int32_t* src, int32_t* dst, int32_t* w) {
  if (scale < 1 || scale > 30 || m < 0 || wmax < 1) return -1;
  return unique_edges(m, wmax, seed, src, dst, w, [&](uint64_t i, uint32_t& u, uint32_t& v) {
    rmat_pair(seed, i, scale, a, b, c, u, v);
  });
}

/// side x side grid, arcs both ways between 4-neighbours, plus ceil(fraction * arcs) distinct
/// uniform shortcut arcs (non-loop, not already present). Returns the edge count (<= cap).
int64_t dyng_gen_grid(int64_t side, double fraction, int32_t wmax, uint64_t seed, int32_t* src, int32_t* dst,
                      int32_t* w, int64_t cap) {
  const int64_t n = side * side;
  const int64_t arcs = 4 * side * (side - 1);
  const int64_t shortcuts = static_cast<int64_t>(std::ceil(fraction * static_cast<double>(arcs)));
  if (arcs + shortcuts > cap) return -1;
  std::vector<uint64_t> keys(static_cast<size_t>(arcs));
  int64_t e = 0;
  for (int64_t r = 0; r < side; ++r) {
    for (int64_t c = 0; c < side; ++c) {
      const uint64_t v = static_cast<uint64_t>(r * side + c);
      if (c + 1 < side) keys[e++] = (v << 32) | (v + 1);
      if (c > 0) keys[e++] = (v << 32) | (v - 1);
      if (r + 1 < side) keys[e++] = (v << 32) | (v + side);
      if (r > 0) keys[e++] = (v << 32) | (v - side);
    }
  }
  std::vector<uint64_t> sorted(keys);
  __gnu_parallel::sort(sorted.begin(), sorted.end());
  std::vector<uint64_t> extra;
  std::unordered_set<uint64_t> seen;
  for (uint64_t i = 0; static_cast<int64_t>(extra.size()) < shortcuts; ++i) {
    const uint64_t u = splitmix(seed ^ 0x5348ULL, 2 * i) % static_cast<uint64_t>(n);
    const uint64_t v = splitmix(seed ^ 0x5348ULL, 2 * i + 1) % static_cast<uint64_t>(n);
    const uint64_t k = (u << 32) | v;
    if (u == v || std::binary_search(sorted.begin(), sorted.end(), k) || !seen.insert(k).second) continue;
    extra.push_back(k);
  }
  keys.insert(keys.end(), extra.begin(), extra.end());
  const int64_t total = static_cast<int64_t>(keys.size());
  std::vector<std::pair<uint64_t, int64_t>> order(static_cast<size_t>(total));
  for (int64_t i = 0; i < total; ++i) order[i] = {keys[i], i};
  __gnu_parallel::sort(order.begin(), order.end());
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < total; ++i) {
    src[i] = static_cast<int32_t>(order[i].first >> 32);
    dst[i] = static_cast<int32_t>(order[i].first & 0xffffffffULL);
    w[i] = 1 + static_cast<int32_t>(splitmix(seed ^ kWeightStream, static_cast<uint64_t>(order[i].second)) %
                                    static_cast<uint64_t>(wmax));
  }
  return total;
}

/// Mixed batches against a graph given as sorted unique (src, dst): per batch n_ins new edges
/// (endpoints uniform, or RMAT when rmat_scale > 0; not in the graph, distinct over all batches),
/// n_del deletions of distinct existing edges and n_rew reweights (upserts of existing edges not
/// deleted by any batch) to U[1, wmax]. Batch b's insertions are its new edges then its reweights.
int dyng_gen_batches(const int32_t* src, const int32_t* dst, int64_t m, int64_t n, int32_t batches, int64_t n_ins,
                     int64_t n_del, int64_t n_rew, int32_t wmax, uint64_t seed, int32_t rmat_scale, double a, double b,
                     double c, int32_t* ins_src, int32_t* ins_dst, int32_t* ins_w, int32_t* del_src, int32_t* del_dst) {
  const int64_t total_del = batches * n_del, total_rew = batches * n_rew, total_ins = batches * n_ins;
  if (total_del + total_rew > m) return -1;
  std::vector<uint64_t> used((static_cast<size_t>(m) + 63) / 64, 0);
  auto pick = [&](uint64_t stream, int64_t count, std::vector<int64_t>& out) {
    out.clear();
    for (uint64_t i = 0; static_cast<int64_t>(out.size()) < count; ++i) {
      const int64_t e = static_cast<int64_t>(splitmix(seed ^ stream, i) % static_cast<uint64_t>(m));
      uint64_t& word = used[static_cast<size_t>(e) / 64];
      const uint64_t bit = 1ULL << (e % 64);
      if (word & bit) continue;
      word |= bit;
      out.push_back(e);
    }
  };
  std::vector<int64_t> dels, rews;
  pick(0xDE1ULL, total_del, dels);
  pick(0x2E3ULL, total_rew, rews);
  auto exists = [&](uint32_t u, uint32_t v) {
    int64_t lo = 0, hi = m;
    const uint64_t k = (static_cast<uint64_t>(u) << 32) | v;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      const uint64_t km = (static_cast<uint64_t>(static_cast<uint32_t>(src[mid])) << 32) | static_cast<uint32_t>(dst[mid]);
      if (km < k) lo = mid + 1; else hi = mid;
    }
    return lo < m && static_cast<uint32_t>(src[lo]) == u && static_cast<uint32_t>(dst[lo]) == v;
  };
  std::unordered_set<uint64_t> fresh;
  fresh.reserve(static_cast<size_t>(total_ins) * 2 + 16);
  std::vector<uint64_t> inserts;
  inserts.reserve(static_cast<size_t>(total_ins));
  const uint64_t ins_seed = seed ^ 0x1A5ULL;
  for (uint64_t i = 0; static_cast<int64_t>(inserts.size()) < total_ins; ++i) {
    uint32_t u, v;
    if (rmat_scale > 0) {
      rmat_pair(ins_seed, i, rmat_scale, a, b, c, u, v);
    } else {
      u = static_cast<uint32_t>(splitmix(ins_seed, 2 * i) % static_cast<uint64_t>(n));
      v = static_cast<uint32_t>(splitmix(ins_seed, 2 * i + 1) % static_cast<uint64_t>(n));
    }
    if (u == v || u >= n || v >= n || exists(u, v)) continue;
    const uint64_t k = (static_cast<uint64_t>(u) << 32) | v;
    if (!fresh.insert(k).second) continue;
    inserts.push_back(k);
  }
  const int64_t per_ins = n_ins + n_rew;
  for (int32_t bt = 0; bt < batches; ++bt) {
    for (int64_t j = 0; j < n_ins; ++j) {
      const uint64_t k = inserts[static_cast<size_t>(bt * n_ins + j)];
      const int64_t o = bt * per_ins + j;
      ins_src[o] = static_cast<int32_t>(k >> 32);
      ins_dst[o] = static_cast<int32_t>(k & 0xffffffffULL);
      ins_w[o] = 1 + static_cast<int32_t>(splitmix(seed ^ 0x1B7ULL, static_cast<uint64_t>(o)) % static_cast<uint64_t>(wmax));
    }
    for (int64_t j = 0; j < n_rew; ++j) {
      const int64_t e = rews[static_cast<size_t>(bt * n_rew + j)];
      const int64_t o = bt * per_ins + n_ins + j;
      ins_src[o] = src[e];
      ins_dst[o] = dst[e];
      ins_w[o] = 1 + static_cast<int32_t>(splitmix(seed ^ 0x2B7ULL, static_cast<uint64_t>(o)) % static_cast<uint64_t>(wmax));
    }
    for (int64_t j = 0; j < n_del; ++j) {
      const int64_t e = dels[static_cast<size_t>(bt * n_del + j)];
      del_src[bt * n_del + j] = src[e];
      del_dst[bt * n_del + j] = dst[e];
    }
  }
  return 0;
}

/// A stream of batches of uneven size (the learned-scheduler configuration): batch b has
/// floor(10^U[log10 lo, log10 hi]) operations, a delete fraction U[0, 1], deletions of distinct
/// edges of the graph (src, dst sorted; an edge is deleted at most once over the stream, and the
/// deletions stop when the graph has none left) and insertions of new RMAT pairs (w ~ U[1, wmax];
/// a pair inserted twice becomes an upsert). Sizes come first (sizes[b] = {insertions,
/// deletions}, a call with null outputs fills only them); then the operations, batch after batch.
int64_t dyng_gen_stream(const int32_t* src, const int32_t* dst, int64_t m, int64_t n, int32_t count, double lo,
                        double hi, int32_t wmax, uint64_t seed, int32_t rmat_scale, double a, double b, double c,
                        int64_t* sizes, int32_t* ins_src, int32_t* ins_dst, int32_t* ins_w, int32_t* del_src,
                        int32_t* del_dst) {
  int64_t left = m, total_ins = 0, total_del = 0;
  for (int32_t bt = 0; bt < count; ++bt) {
    const double e = std::log10(lo) + unit(splitmix(seed ^ 0x53495AULL, 2 * bt)) * (std::log10(hi) - std::log10(lo));
    const int64_t size = std::max<int64_t>(1, static_cast<int64_t>(std::pow(10.0, e)));
    int64_t nd = static_cast<int64_t>(static_cast<double>(size) * unit(splitmix(seed ^ 0x53495AULL, 2 * bt + 1)));
    nd = std::min(nd, left);
    left -= nd;
    sizes[2 * bt] = size - nd;
    sizes[2 * bt + 1] = nd;
    total_ins += size - nd;
    total_del += nd;
  }
  if (ins_src == nullptr) return total_ins + total_del;
  std::vector<uint64_t> used((static_cast<size_t>(m) + 63) / 64, 0);
  int64_t di = 0, ii = 0;
  uint64_t draw = 0, idraw = 0;
  const uint64_t ins_seed = seed ^ 0x5157ULL;
  for (int32_t bt = 0; bt < count; ++bt) {
    for (int64_t j = 0; j < sizes[2 * bt + 1]; ++j) {
      for (;;) {
        const int64_t e = static_cast<int64_t>(splitmix(seed ^ 0xDE15ULL, draw++) % static_cast<uint64_t>(m));
        uint64_t& word = used[static_cast<size_t>(e) / 64];
        const uint64_t bit = 1ULL << (e % 64);
        if (word & bit) continue;
        word |= bit;
        del_src[di] = src[e];
        del_dst[di] = dst[e];
        ++di;
        break;
      }
    }
    for (int64_t j = 0; j < sizes[2 * bt]; ++j) {
      uint32_t u, v;
      do {
        rmat_pair(ins_seed, idraw++, rmat_scale, a, b, c, u, v);
      } while (u == v || u >= n || v >= n);
      ins_src[ii] = static_cast<int32_t>(u);
      ins_dst[ii] = static_cast<int32_t>(v);
      ins_w[ii] = 1 + static_cast<int32_t>(splitmix(seed ^ 0x1B7ULL, static_cast<uint64_t>(ii)) % static_cast<uint64_t>(wmax));
      ++ii;
    }
  }
  return total_ins + total_del;
}

/// A batch aimed at the current shortest-path tree (the high-diameter grid configuration):
/// `n_tree` deletions of distinct tree edges (parent[v], v) for random reachable v, `n_rand`
/// deletions of random grid arcs (a grid arc deleted before makes an ignored deletion), and
/// `n_ins` insertions of random non-loop arcs with w ~ U[1, wmax] (an arc already present is an
/// upsert). Every draw is splitmix64 of (seed, batch, draw index), so the batch depends only on
/// the tree it is drawn against. Returns 0, or -1 when the tree has too few edges.
int dyng_gen_tree_batch(int64_t side, const int32_t* parents, int64_t n_ins, int64_t n_tree, int64_t n_rand,
                        int32_t wmax, uint64_t seed, int64_t batch, int32_t* ins_src, int32_t* ins_dst,
                        int32_t* ins_w, int32_t* del_src, int32_t* del_dst) {
  const int64_t n = side * side;
  const uint64_t s = seed ^ (0x7452ULL << 40) ^ static_cast<uint64_t>(batch) * 0x9E3779B97F4A7C15ULL;
  int64_t with_parent = 0;
  for (int64_t v = 0; v < n && with_parent <= n_tree; ++v) with_parent += parents[v] >= 0 ? 1 : 0;
  if (with_parent < n_tree) return -1;
  std::vector<uint8_t> taken(static_cast<size_t>(n), 0);
  int64_t d = 0;
  for (uint64_t i = 0; d < n_tree; ++i) {
    const int64_t v = static_cast<int64_t>(splitmix(s, 3 * i) % static_cast<uint64_t>(n));
    if (parents[v] < 0 || taken[v]) continue;
    taken[v] = 1;
    del_src[d] = parents[v];
    del_dst[d] = static_cast<int32_t>(v);
    ++d;
  }
  for (int64_t i = 0; i < n_rand; ++i) {
    const int64_t u = static_cast<int64_t>(splitmix(s, 3 * i + 1) % static_cast<uint64_t>(n));
    const int64_t r = u / side, c = u % side;
    int64_t v = -1;
    for (uint64_t k = 0; v < 0; ++k) {
      switch (splitmix(s ^ 0x444952ULL, 4 * static_cast<uint64_t>(i) + k) % 4) {
        case 0: if (c + 1 < side) v = u + 1; break;
        case 1: if (c > 0) v = u - 1; break;
        case 2: if (r + 1 < side) v = u + side; break;
        default: if (r > 0) v = u - side; break;
      }
    }
    del_src[d] = static_cast<int32_t>(u);
    del_dst[d] = static_cast<int32_t>(v);
    ++d;
  }
  for (int64_t i = 0; i < n_ins; ++i) {
    uint64_t u, v;
    uint64_t k = 0;
    do {
      u = splitmix(s ^ 0x494E53ULL, 2 * (static_cast<uint64_t>(i) * 8 + k)) % static_cast<uint64_t>(n);
      v = splitmix(s ^ 0x494E53ULL, 2 * (static_cast<uint64_t>(i) * 8 + k) + 1) % static_cast<uint64_t>(n);
      ++k;
    } while (u == v);
    ins_src[i] = static_cast<int32_t>(u);
    ins_dst[i] = static_cast<int32_t>(v);
    ins_w[i] = 1 + static_cast<int32_t>(splitmix(s ^ kWeightStream, static_cast<uint64_t>(i)) % static_cast<uint64_t>(wmax));
  }
  return 0;
}

}  // extern "C"
