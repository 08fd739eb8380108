// MOSP's combined solve on the device (cpp/src/algorithms/mosp/mosp.cpp:264-330, the caller of
// sssp_solve_view): the combined graph of K shortest-path trees (MOSP-OpenMP's
// combinedGraphSospCpu, restated by cpp/src/algorithms/mosp/openmp.cpp:39-118) written straight
// into out-rows, and the path costs of the MOSP tree (cpp/src/algorithms/mosp/sequential.cpp:
// 172-236) by path doubling instead of a traversal from the source. The solve itself is the
// persistent engine of sssp.hip on the combined graph.
//
// The combined graph is a solve-only graph: out-rows in one arena (the edge list sorted by parent,
// row u its run; the order within a row is that of the list, which the engine's canonical tree
// does not depend on), no in-rows (a solve from the source never pulls), no slack. Its buffers live in a per-device cache
// and are reused by the next call (they grow, never shrink, until the process ends).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <vector>

namespace dyng_hip {

namespace {

constexpr int kBlock = 1024;
constexpr int kWaves = kBlock / 64;

struct combine_params {
  const int32_t* const* parents;  ///< K device pointers to the objectives' parent arrays
  const int32_t* terms;           ///< K preference terms L / pref[k]
  int32_t K;
  int32_t n;
  int32_t source;
  int32_t base;  ///< L (K + 1)
};

/// The combined edge that objective k's tree gives v, if any (MOSP's combinedEdge,
/// openmp.cpp:39-63): the tree edge (p, v) counted once, with its first objective, weighing
/// base less the terms of every objective whose tree holds it.
DYNG_DEV bool combined_edge(const combine_params& c, int32_t v, int k, int32_t& p, int32_t& w) {
  p = c.parents[k][v];
  if (p < 0) return false;
  for (int j = 0; j < k; ++j) {
    if (c.parents[j][v] == p) return false;
  }
  int32_t weight = c.base - c.terms[k];
  for (int j = k + 1; j < c.K; ++j) {
    if (c.parents[j][v] == p) weight -= c.terms[j];
  }
  w = weight;
  return true;
}

/// One thread per vertex v: its combined edges appended to a flat list (parent key, then
/// (v, weight) as one word) with one atomic per workgroup (a workgroup scan of the threads' edge
/// counts), plus the weight sum.
__global__ void __launch_bounds__(kBlock) k_combine(combine_params c, uint32_t* keys, u64* vals,
                                                    unsigned long long* totals) {
  __shared__ int wave_total[kWaves];
  __shared__ int block_base;
  __shared__ unsigned long long part[kWaves];
  const int32_t v = static_cast<int32_t>(blockIdx.x * kBlock + threadIdx.x);
  const bool live = v < c.n && v != c.source;
  int mine = 0;
  unsigned long long wsum = 0;
  if (live) {
    for (int k = 0; k < c.K; ++k) {
      int32_t p, x;
      if (combined_edge(c, v, k, p, x)) {
        ++mine;
        wsum += static_cast<unsigned long long>(x);
      }
    }
  }
  int incl = mine;  // inclusive scan of `mine` over the wave
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off);
    if (lane() >= off) incl += y;
  }
  const int w_id = static_cast<int>(threadIdx.x) / 64;
  if (lane() == 63) wave_total[w_id] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < kWaves; ++i) {
      const int t = wave_total[i];
      wave_total[i] = run;
      run += t;
    }
    block_base = run > 0 ? static_cast<int>(atomicAdd(&totals[0], static_cast<unsigned long long>(run))) : 0;
  }
  __syncthreads();
  int slot = block_base + wave_total[w_id] + incl - mine;
  if (mine > 0) {
    for (int k = 0; k < c.K; ++k) {
      int32_t p, x;
      if (combined_edge(c, v, k, p, x)) {
        keys[slot] = static_cast<uint32_t>(p);
        vals[slot] = (static_cast<u64>(static_cast<uint32_t>(v)) << 32) | static_cast<uint32_t>(x);
        ++slot;
      }
    }
  }
  wsum = block_reduce(wsum, part, op_sum{});
  if (threadIdx.x == 0 && wsum) atomicAdd(&totals[1], wsum);
}

/// One thread per edge of the list sorted by parent: its (neighbour, weight) into the row arena
/// at its index, and at the first edge of each parent's run that row's bounds {offset, length}
/// (the run's end found by a galloping search). Rows of vertices without edges were cleared.
__global__ void k_combined_rows(const uint32_t* keys, const u64* vals, int64_t m, int32_t* nbr, int32_t* w,
                                int2* meta, int32_t* base_len) {
  const int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x;
  if (i >= m) return;
  const u64 e = vals[i];
  nbr[i] = static_cast<int32_t>(e >> 32);
  w[i] = static_cast<int32_t>(static_cast<uint32_t>(e));
  const uint32_t u = keys[i];
  if (i > 0 && keys[i - 1] == u) return;
  // the run of u is [i, end): gallop to a bracket, then bisect
  int64_t lo = i, step = 1;  // keys[lo] == u
  int64_t hi = i + 1;
  while (hi < m && keys[hi] == u) {
    lo = hi;
    step <<= 1;
    hi = lo + step < m ? lo + step : m;
  }
  while (hi - lo > 1) {  // keys[lo] == u, and hi == m or keys[hi] != u
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] == u) lo = mid; else hi = mid;
  }
  const int len = static_cast<int>(hi - i);
  meta[u] = make_int2(static_cast<int>(i), len);
  base_len[u] = len;
}

template <typename T>
__global__ void k_fill_value(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < n; i += int64_t{gridDim.x} * blockDim.x) p[i] = v;
}

/// The slot of the first alive entry with neighbour `x` in a row of two sorted runs ([0, bl) then
/// [bl, len)), or -1.
DYNG_DEV int first_alive(const int32_t* row, int bl, int len, int32_t x) {
  for (int r = 0; r < 2; ++r) {
    const int lo = r == 0 ? 0 : bl, hi = r == 0 ? bl : len;
    for (int j = row_lower_bound(row, lo, hi, x); j < hi && (row[j] & 0x7fffffff) == x; ++j) {
      if (row[j] >= 0) return j;
    }
  }
  return -1;
}

/// Path costs, first step: every vertex's own tree edge (p, v) as the first alive (p, v) of p's
/// out-row (the base run, then the tail: MOSP's row order), its first `columns` weights; the
/// source starts at 0 and an unreachable vertex at INF, both with no ancestor. Without parallel
/// edges the pair (p, v) is one edge, found in v's in-row instead (short and read in vertex
/// order, where p's out-row may be a hub's). A tree edge that is not an edge of the graph is
/// reported (the lowest such vertex). active[0] is set when some vertex has an ancestor.
struct edge_rows {
  const int2* meta;
  const int32_t* base;
  const int32_t* nbr;
  const int32_t* w;
  int64_t slot_cap;
};

__global__ void k_tree_edge_costs(edge_rows out, edge_rows in, bool by_in_row, int columns, const int32_t* parents,
                                  int32_t n, int32_t source, int64_t* acc, int32_t* jump, int* missing, int* active) {
  const int32_t v = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  int32_t link = -1;
  if (v < n) {
    const int32_t p = v == source ? -1 : parents[v];
    int64_t* a = acc + static_cast<int64_t>(v) * columns;
    int64_t slot = -1;
    const int32_t* w = nullptr;
    int64_t cap = 0;
    if (p >= 0) {
      const edge_rows& er = by_in_row ? in : out;
      const int32_t row_of = by_in_row ? v : p;
      const int32_t want = by_in_row ? p : v;
      const int2 mt = er.meta[row_of];
      const int hit = first_alive(er.nbr + mt.x, er.base[row_of], mt.y, want);
      if (hit >= 0) {
        slot = mt.x + int64_t{hit};
        w = er.w;
        cap = er.slot_cap;
      } else {
        atomicMin(missing, v);
      }
    }
    if (slot >= 0) {
      for (int c = 0; c < columns; ++c) a[c] = w[c * cap + slot];
      link = p;
    } else {
      const int64_t start = v == source ? 0 : sssp_infinity;
      for (int c = 0; c < columns; ++c) a[c] = start;
    }
    jump[v] = link;
  }
  if (__syncthreads_or(link >= 0) && threadIdx.x == 0) active[0] = 1;
}

/// Path doubling, round r: each vertex adds its ancestor's sum and jumps to the ancestor's
/// ancestor. A round whose input has no vertex left with an ancestor does nothing; a round that
/// runs records itself in `last` and sets active[r + 1] when some vertex still has an ancestor.
/// A vertex's jump is -1 in the round after its sum became final (that round copies the sum to
/// the other buffer) and -2 from then on (both buffers hold it: the round only passes the -2 on).
__global__ void k_path_double(const int64_t* acc_in, const int32_t* jump_in, int64_t* acc_out, int32_t* jump_out,
                              int32_t n, int columns, int round, int* active, int* last) {
  if (active[round] == 0) return;  // uniform: every thread reads the same flag
  bool more = false;  // some vertex of this thread still has an ancestor
  const int64_t stride = int64_t{gridDim.x} * blockDim.x;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < n; i += stride) {
    const int32_t v = static_cast<int32_t>(i);
    int32_t next = -2;
    const int32_t j = jump_in[v];
    if (j == -1) {
      const int64_t* a = acc_in + static_cast<int64_t>(v) * columns;
      int64_t* o = acc_out + static_cast<int64_t>(v) * columns;
      for (int c = 0; c < columns; ++c) o[c] = a[c];
    } else if (j >= 0) {
      const int64_t* a = acc_in + static_cast<int64_t>(v) * columns;
      int64_t* o = acc_out + static_cast<int64_t>(v) * columns;
      const int64_t* b = acc_in + static_cast<int64_t>(j) * columns;
      for (int c = 0; c < columns; ++c) o[c] = a[c] + b[c];
      const int32_t up = jump_in[j];
      next = up < 0 ? -1 : up;
    }
    jump_out[v] = next;
    more = more || next >= 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *last = round;
  if (__syncthreads_or(more) && threadIdx.x == 0) active[round + 1] = 1;
}

/// Round r writes buffer (r + 1) % 2 (buffer 0 is the caller's, which the first step wrote): the
/// sums are in `acc` only when the last round that ran is even; then copy them home.
__global__ void k_costs_home(const int64_t* acc, int64_t* costs, int64_t count, const int* last) {
  const int r = *last;
  if (r < 0 || (r & 1) == 1) return;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count; i += int64_t{gridDim.x} * blockDim.x) {
    costs[i] = acc[i];
  }
}

/// Counts of the threads of a workgroup turned into slots of a shared list: the workgroup's
/// exclusive scan of `mine`, then one atomic per workgroup on `counter`. Every thread calls it;
/// returns this thread's first slot. `wave_total` and `block_base` are LDS scratch.
DYNG_DEV int block_slots(int mine, unsigned long long* counter, int* wave_total, int* block_base) {
  int incl = mine;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(incl, off);
    if (lane() >= off) incl += y;
  }
  const int w_id = static_cast<int>(threadIdx.x) / 64;
  __syncthreads();
  if (lane() == 63) wave_total[w_id] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < kWaves; ++i) {
      const int t = wave_total[i];
      wave_total[i] = run;
      run += t;
    }
    *block_base = run > 0 ? static_cast<int>(atomicAdd(counter, static_cast<unsigned long long>(run))) : 0;
  }
  __syncthreads();
  return *block_base + wave_total[w_id] + incl - mine;
}

/// The combined graph as an edge list (src, dst, weight), one thread per vertex, with its weight
/// sum: the input of the kept combined graph's first build.
__global__ void __launch_bounds__(kBlock) k_combine_list(combine_params c, int32_t* src, int32_t* dst, int32_t* w,
                                                         unsigned long long* totals) {
  __shared__ int wave_total[kWaves];
  __shared__ int block_base;
  __shared__ unsigned long long part[kWaves];
  const int32_t v = static_cast<int32_t>(blockIdx.x * kBlock + threadIdx.x);
  const bool live = v < c.n && v != c.source;
  int mine = 0;
  unsigned long long wsum = 0;
  if (live) {
    for (int k = 0; k < c.K; ++k) {
      int32_t p, x;
      if (combined_edge(c, v, k, p, x)) {
        ++mine;
        wsum += static_cast<unsigned long long>(x);
      }
    }
  }
  int slot = block_slots(mine, &totals[0], wave_total, &block_base);
  if (mine > 0) {
    for (int k = 0; k < c.K; ++k) {
      int32_t p, x;
      if (combined_edge(c, v, k, p, x)) {
        src[slot] = p;
        dst[slot] = v;
        w[slot] = x;
        ++slot;
      }
    }
  }
  wsum = block_reduce(wsum, part, op_sum{});
  if (threadIdx.x == 0 && wsum) atomicAdd(&totals[1], wsum);
}

/// Up to kRegTrees trees: a vertex's K parents and the K terms held in registers (static
/// indices only), so each parent is read once instead of once per comparison.
constexpr int kRegTrees = 8;

struct reg_trees {
  int32_t p[kRegTrees];
  int32_t t[kRegTrees];
};

DYNG_DEV void load_trees(const combine_params& c, int32_t v, bool live, reg_trees& r) {
#pragma unroll
  for (int k = 0; k < kRegTrees; ++k) {
    r.p[k] = (live && k < c.K) ? c.parents[k][v] : -1;
    r.t[k] = k < c.K ? c.terms[k] : 0;
  }
}

/// combined_edge over the registers: objective k's edge (p, v) if it is the first objective with
/// that parent, weighing base less the terms of every objective holding it.
DYNG_DEV bool reg_edge(const reg_trees& r, int32_t base, int k, int32_t& p, int32_t& w) {
  p = -1;
  w = 0;
#pragma unroll
  for (int i = 0; i < kRegTrees; ++i) {
    if (i == k) p = r.p[i];
  }
  if (p < 0) return false;
  int32_t weight = base;
#pragma unroll
  for (int j = 0; j < kRegTrees; ++j) {
    if (r.p[j] == p) {
      if (j < k) return false;
      weight -= r.t[j];
    }
  }
  w = weight;
  return true;
}

/// k_combine_diff for at most kRegTrees trees, from registers.
__global__ void __launch_bounds__(kBlock) k_combine_diff_reg(combine_params old_c, combine_params new_c,
                                                             int32_t* del_src, int32_t* del_dst, int32_t* ins_src,
                                                             int32_t* ins_dst, int32_t* ins_w,
                                                             unsigned long long* totals) {
  __shared__ int wave_total[kWaves];
  __shared__ int block_base;
  const int32_t v = static_cast<int32_t>(blockIdx.x * kBlock + threadIdx.x);
  const bool has_old = v < old_c.n && v != old_c.source;
  const bool has_new = v < new_c.n && v != new_c.source;
  reg_trees ro, rn;
  load_trees(old_c, v, has_old, ro);
  load_trees(new_c, v, has_new, rn);
  // an old edge whose parent no new tree gives v is deleted; a new edge is inserted unless the old
  // trees gave the same (p, v) the same weight
  unsigned del_mask = 0, ins_mask = 0;
#pragma unroll
  for (int k = 0; k < kRegTrees; ++k) {
    int32_t p, x;
    if (reg_edge(ro, old_c.base, k, p, x)) {
      bool kept = false;
#pragma unroll
      for (int j = 0; j < kRegTrees; ++j) kept = kept || rn.p[j] == p;
      if (!kept) del_mask |= 1u << k;
    }
    if (reg_edge(rn, new_c.base, k, p, x)) {
      bool same = false;
#pragma unroll
      for (int j = 0; j < kRegTrees; ++j) {
        int32_t po, xo;
        same = same || (reg_edge(ro, old_c.base, j, po, xo) && po == p && xo == x);
      }
      if (!same) ins_mask |= 1u << k;
    }
  }
  int del_at = block_slots(__popc(del_mask), &totals[2], wave_total, &block_base);
  int ins_at = block_slots(__popc(ins_mask), &totals[3], wave_total, &block_base);
#pragma unroll
  for (int k = 0; k < kRegTrees; ++k) {
    int32_t p, x;
    if ((del_mask >> k) & 1u) {
      reg_edge(ro, old_c.base, k, p, x);
      del_src[del_at] = p;
      del_dst[del_at] = v;
      ++del_at;
    }
    if ((ins_mask >> k) & 1u) {
      reg_edge(rn, new_c.base, k, p, x);
      ins_src[ins_at] = p;
      ins_dst[ins_at] = v;
      ins_w[ins_at] = x;
      ++ins_at;
    }
  }
}

/// Whether any of the K trees gives v the parent p.
DYNG_DEV bool some_tree_has(const combine_params& c, int32_t v, int32_t p) {
  for (int j = 0; j < c.K; ++j) {
    if (c.parents[j][v] == p) return true;
  }
  return false;
}

/// The batch that turns the combined graph of the old trees into that of the new ones, one
/// thread per vertex v: a combined in-edge (p, v) of the old trees that the new ones lack is
/// deleted, one the new trees have is inserted (an upsert) unless the old trees gave it the same
/// weight. Vertices past old.n had no combined edges. Each thread counts its operations, the
/// workgroup claims their slots (totals[2], totals[3]: deletions, insertions), then the thread
/// writes them.
__global__ void __launch_bounds__(kBlock) k_combine_diff(combine_params old_c, combine_params new_c,
                                                         int32_t* del_src, int32_t* del_dst, int32_t* ins_src,
                                                         int32_t* ins_dst, int32_t* ins_w,
                                                         unsigned long long* totals) {
  __shared__ int wave_total[kWaves];
  __shared__ int block_base;
  const int32_t v = static_cast<int32_t>(blockIdx.x * kBlock + threadIdx.x);
  const bool has_old = v < old_c.n && v != old_c.source;
  const bool has_new = v < new_c.n && v != new_c.source;
  int dels = 0, ins = 0;
  auto each = [&](bool write, int del_at, int ins_at) {
    if (has_old) {
      for (int k = 0; k < old_c.K; ++k) {
        int32_t p, x;
        if (combined_edge(old_c, v, k, p, x) && !(has_new && some_tree_has(new_c, v, p))) {
          if (write) {
            del_src[del_at] = p;
            del_dst[del_at] = v;
          }
          ++del_at;
        }
      }
    }
    if (has_new) {
      for (int k = 0; k < new_c.K; ++k) {
        int32_t p, x;
        if (!combined_edge(new_c, v, k, p, x)) continue;
        bool same = false;
        if (has_old) {
          for (int j = 0; j < old_c.K && !same; ++j) {
            int32_t po, xo;
            same = combined_edge(old_c, v, j, po, xo) && po == p && xo == x;
          }
        }
        if (!same) {
          if (write) {
            ins_src[ins_at] = p;
            ins_dst[ins_at] = v;
            ins_w[ins_at] = x;
          }
          ++ins_at;
        }
      }
    }
    dels = del_at;
    ins = ins_at;
  };
  each(false, 0, 0);
  const int nd = dels, ni = ins;
  const int del_at = block_slots(nd, &totals[2], wave_total, &block_base);
  const int ins_at = block_slots(ni, &totals[3], wave_total, &block_base);
  if (nd > 0 || ni > 0) each(true, del_at, ins_at);
}

unsigned int grid(int64_t work, int block) {
  return static_cast<unsigned int>(std::max<int64_t>(1, (work + block - 1) / block));
}

/// The solve-only combined graph and the scratch of its build, one per device.
struct combined_cache {
  std::unique_ptr<dyng_hip_graph> g;
  dbuf<uint32_t> keys_a, keys_b;
  dbuf<u64> vals_a, vals_b;
  dbuf<unsigned char> sort_tmp;
  dbuf<const int32_t*> parents;
  dbuf<int32_t> terms;
  dbuf<unsigned long long> totals;
  dbuf<int64_t> acc;
  dbuf<int32_t> jump_a, jump_b;
  dbuf<int> flags;  ///< missing vertex, last round, then the rounds' active flags
  const int2* in_meta_cleared = nullptr;  ///< the in-row meta array last cleared
};

combined_cache& cache_for(int device) {
  static std::vector<std::unique_ptr<combined_cache>> caches;
  if (device < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: bad device");
  if (static_cast<size_t>(device) >= caches.size()) caches.resize(static_cast<size_t>(device) + 1);
  if (!caches[device]) caches[device] = std::make_unique<combined_cache>();
  return *caches[device];
}

}  // namespace

dyng_hip_graph* mosp_build_combined(int device, const int32_t* const* parents, int32_t num_objectives, int64_t n,
                                    int32_t source, int32_t base, const int32_t* terms, hipStream_t s,
                                    int64_t* edges, int64_t* weight_sum) {
  combined_cache& cc = cache_for(device);
  if (!cc.g) cc.g = std::make_unique<dyng_hip_graph>();
  dyng_hip_graph* g = cc.g.get();
  cc.parents.reserve(num_objectives);
  cc.terms.reserve(num_objectives);
  cc.totals.reserve(2);
  const int64_t cap = std::max<int64_t>(1, n * num_objectives);
  cc.keys_a.reserve(cap);
  cc.keys_b.reserve(cap);
  cc.vals_a.reserve(cap);
  cc.vals_b.reserve(cap);
  // small synchronous copies from the caller's host arrays (the previous call's work on the cache
  // ended before it returned)
  DYNG_HIP_TRY(hipMemcpy(cc.parents.p, parents, num_objectives * sizeof(const int32_t*), hipMemcpyHostToDevice));
  DYNG_HIP_TRY(hipMemcpy(cc.terms.p, terms, num_objectives * sizeof(int32_t), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_fill_value<unsigned long long>, dim3(1), dim3(64), 0, s, cc.totals.p, int64_t{2}, 0ULL);
  combine_params c{cc.parents.p, cc.terms.p, num_objectives, static_cast<int32_t>(n), source, base};
  hipLaunchKernelGGL(k_combine, dim3(grid(n, kBlock)), dim3(kBlock), 0, s, c, cc.keys_a.p, cc.vals_a.p, cc.totals.p);
  DYNG_HIP_TRY(hipGetLastError());
  unsigned long long host[2] = {0, 0};
  DYNG_HIP_TRY(hipMemcpyAsync(host, cc.totals.p, sizeof(host), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  const int64_t m = static_cast<int64_t>(host[0]);
  if (m >= INT32_MAX) fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the combined graph has 2^31 or more edges");
  *edges = m;
  *weight_sum = static_cast<int64_t>(host[1]);
  // group the edges by parent (a stable radix sort over the bits of a vertex id)
  int bits = 1;
  while ((int64_t{1} << bits) < n) ++bits;
  if (m > 0) {
    size_t bytes = 0;
    DYNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, cc.keys_a.p, cc.keys_b.p, cc.vals_a.p, cc.vals_b.p,
                                           static_cast<size_t>(m), 0, bits, s));
    cc.sort_tmp.reserve(bytes);
    DYNG_HIP_TRY(rocprim::radix_sort_pairs(static_cast<void*>(cc.sort_tmp.p), bytes, cc.keys_a.p, cc.keys_b.p,
                                           cc.vals_a.p, cc.vals_b.p, static_cast<size_t>(m), 0, bits, s));
  }
  // the graph: out-rows in one arena of exactly m slots, each row where its run lies in the
  // sorted list
  g->device = device;
  g->props = dyng_hip_graph_props{1, 1, 0, 0, 0, 0, 1, 1, 0.0};
  g->K = 1;
  g->n = n;
  g->m = m;
  g->n_cap = std::max<int64_t>(g->n_cap, n);
  g->out.meta.reserve(g->n_cap);
  g->out.base.reserve(g->n_cap);
  g->out.nbr.reserve(std::max<int64_t>(m, 1));
  g->out.w.reserve(std::max<int64_t>(m, 1));
  g->out.slot_cap = std::max<int64_t>(m, 1);
  g->out.tail = m;
  g->out.garbage = 0;
  // no in-rows: every vertex has an empty in-row
  g->in.meta.reserve(g->n_cap);
  if (cc.in_meta_cleared != g->in.meta.p) {  // a new allocation
    hipLaunchKernelGGL(k_fill_value<int2>, dim3(std::min(grid(g->n_cap, 256), 4096u)), dim3(256), 0, s,
                       g->in.meta.p, g->n_cap, make_int2(0, 0));
    cc.in_meta_cleared = g->in.meta.p;
  }
  g->in.nbr.reserve(1);
  g->in.w.reserve(1);
  g->in.slot_cap = 0;
  g->wmax.assign(1, std::max<int64_t>(base, 1));
  g->wsum.assign(1, *weight_sum);
  g->wmax_count.assign(1, 0);
  g->wmin.assign(1, 1);
  g->last.valid = false;
  g->num_del_slots = 0;
  g->num_ins_slots = 0;
  hipLaunchKernelGGL(k_fill_value<int2>, dim3(std::min(grid(n, 256), 4096u)), dim3(256), 0, s, g->out.meta.p, n,
                     make_int2(0, 0));
  hipLaunchKernelGGL(k_fill_value<int32_t>, dim3(std::min(grid(n, 256), 4096u)), dim3(256), 0, s, g->out.base.p, n,
                     0);
  if (m > 0) {
    hipLaunchKernelGGL(k_combined_rows, dim3(grid(m, 256)), dim3(256), 0, s, cc.keys_b.p, cc.vals_b.p, m,
                       g->out.nbr.p, g->out.w.p, g->out.meta.p, g->out.base.p);
  }
  DYNG_HIP_TRY(hipGetLastError());
  return g;
}

int64_t mosp_combined_list(const int32_t* const* parents_dev, int32_t num_objectives, int64_t n, int32_t source,
                           int32_t base, const int32_t* terms_dev, int32_t* src, int32_t* dst, int32_t* w,
                           unsigned long long* totals, hipStream_t s, int64_t* weight_sum) {
  hipLaunchKernelGGL(k_fill_value<unsigned long long>, dim3(1), dim3(64), 0, s, totals, int64_t{2}, 0ULL);
  combine_params c{parents_dev, terms_dev, num_objectives, static_cast<int32_t>(n), source, base};
  hipLaunchKernelGGL(k_combine_list, dim3(grid(n, kBlock)), dim3(kBlock), 0, s, c, src, dst, w, totals);
  DYNG_HIP_TRY(hipGetLastError());
  unsigned long long host[2] = {0, 0};
  DYNG_HIP_TRY(hipMemcpyAsync(host, totals, sizeof(host), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  *weight_sum = static_cast<int64_t>(host[1]);
  return static_cast<int64_t>(host[0]);
}

void mosp_combined_diff(const int32_t* const* old_parents, int64_t old_n, const int32_t* const* new_parents,
                        int64_t new_n, int32_t num_objectives, int32_t source, int32_t base, const int32_t* terms_dev,
                        int32_t* del_src, int32_t* del_dst, int32_t* ins_src, int32_t* ins_dst, int32_t* ins_w,
                        unsigned long long* totals, hipStream_t s, int64_t* num_del, int64_t* num_ins) {
  hipLaunchKernelGGL(k_fill_value<unsigned long long>, dim3(1), dim3(64), 0, s, totals, int64_t{4}, 0ULL);
  combine_params oc{old_parents, terms_dev, num_objectives, static_cast<int32_t>(old_n), source, base};
  combine_params nc{new_parents, terms_dev, num_objectives, static_cast<int32_t>(new_n), source, base};
  const int64_t n = std::max(old_n, new_n);
  if (num_objectives <= kRegTrees) {
    hipLaunchKernelGGL(k_combine_diff_reg, dim3(grid(n, kBlock)), dim3(kBlock), 0, s, oc, nc, del_src, del_dst,
                       ins_src, ins_dst, ins_w, totals);
  } else {
    hipLaunchKernelGGL(k_combine_diff, dim3(grid(n, kBlock)), dim3(kBlock), 0, s, oc, nc, del_src, del_dst, ins_src,
                       ins_dst, ins_w, totals);
  }
  DYNG_HIP_TRY(hipGetLastError());
  unsigned long long host[4] = {0, 0, 0, 0};
  DYNG_HIP_TRY(hipMemcpyAsync(host, totals, sizeof(host), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  *num_del = static_cast<int64_t>(host[2]);
  *num_ins = static_cast<int64_t>(host[3]);
}

void mosp_path_costs(const dyng_hip_graph* g, const int32_t* parents, int64_t n, int32_t source, int32_t columns,
                     int64_t* costs, hipStream_t s) {
  if (n == 0 || columns == 0) return;
  combined_cache& cc = cache_for(g->device);
  int rounds = 1;
  while ((int64_t{1} << rounds) < n) ++rounds;
  rounds += 1;
  cc.acc.reserve(static_cast<size_t>(n) * columns);
  cc.jump_a.reserve(n);
  cc.jump_b.reserve(n);
  cc.flags.reserve(rounds + 3);
  // flags: [0] the lowest vertex whose tree edge is missing, [1] the last round that ran, then
  // active[0 .. rounds]
  std::vector<int> init(static_cast<size_t>(rounds) + 3, 0);
  init[0] = INT32_MAX;
  init[1] = -1;
  DYNG_HIP_TRY(hipMemcpyAsync(cc.flags.p, init.data(), init.size() * sizeof(int), hipMemcpyHostToDevice, s));
  int* missing = cc.flags.p;
  int* last = cc.flags.p + 1;
  int* active = cc.flags.p + 2;
  // the first step writes the caller's array; the rounds then alternate between it and `acc`
  int64_t* bufs[2] = {costs, cc.acc.p};
  int32_t* jumps[2] = {cc.jump_a.p, cc.jump_b.p};
  const edge_rows out{g->out.meta.p, g->out.base.p, g->out.nbr.p, g->out.w.p, g->out.slot_cap};
  const edge_rows in{g->in.meta.p, g->in.base.p, g->in.nbr.p, g->in.w.p, g->in.slot_cap};
  const bool by_in_row = g->props.allow_parallel == 0;
  hipLaunchKernelGGL(k_tree_edge_costs, dim3(grid(n, 256)), dim3(256), 0, s, out, in, by_in_row, columns, parents,
                     static_cast<int32_t>(n), source, bufs[0], jumps[0], missing, active);
  int cur = 0;
  for (int r = 0; r < rounds; ++r) {
    // a bounded grid: a round after the last one still launches, and returns at once
    hipLaunchKernelGGL(k_path_double, dim3(std::min(grid(n, 256), 2048u)), dim3(256), 0, s, bufs[cur], jumps[cur],
                       bufs[cur ^ 1], jumps[cur ^ 1], static_cast<int32_t>(n), columns, r, active, last);
    cur ^= 1;
  }
  hipLaunchKernelGGL(k_costs_home, dim3(std::min(grid(n * columns, 256), 8192u)), dim3(256), 0, s, cc.acc.p, costs,
                     n * columns, last);
  DYNG_HIP_TRY(hipGetLastError());
  int host_missing = INT32_MAX;
  DYNG_HIP_TRY(hipMemcpyAsync(&host_missing, missing, sizeof(int), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (host_missing != INT32_MAX) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the MOSP tree edge (" + std::to_string(host_missing) +
                                        "'s parent, " + std::to_string(host_missing) +
                                        ") is not an edge of the graph (do the trees belong to this graph?)");
  }
}

}  // namespace dyng_hip
