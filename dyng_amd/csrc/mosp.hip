// MOSP's combined solve on the device (cpp/src/algorithms/mosp/mosp.cpp:264-330, the caller of
// sssp_solve_view): the combined graph of K shortest-path trees, built as a device edge list
// (one thread per vertex, its K tree edges; MOSP-OpenMP's combinedGraphSospCpu restated as
// cpp/src/algorithms/mosp/openmp.cpp:39-118 does), and the path costs of the MOSP tree
// (cpp/src/algorithms/mosp/sequential.cpp:172-236) by path doubling instead of a traversal from
// the source. The solve itself is the persistent engine of sssp.hip on the combined graph.
#include "internal.hpp"

#include <hip/hip_runtime.h>

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

/// One thread per vertex: its combined edges, appended to the edge list with one atomic per
/// workgroup (a workgroup scan of the threads' edge counts), plus the edges' weight sum.
__global__ void __launch_bounds__(kBlock) k_combine(combine_params c, int32_t* src, int32_t* dst, int32_t* w,
                                                    unsigned long long* count, unsigned long long* weight_sum) {
  __shared__ int wave_total[kWaves];
  __shared__ int block_base;
  __shared__ unsigned long long part[kWaves];
  const int32_t v = static_cast<int32_t>(blockIdx.x * kBlock + threadIdx.x);
  int mine = 0;
  unsigned long long wsum = 0;
  if (v < c.n && v != c.source) {
    for (int k = 0; k < c.K; ++k) {
      int32_t p, x;
      if (combined_edge(c, v, k, p, x)) {
        ++mine;
        wsum += static_cast<unsigned long long>(x);
      }
    }
  }
  // exclusive scan of `mine` over the workgroup
  int incl = mine;
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
    block_base = run > 0 ? static_cast<int>(atomicAdd(count, static_cast<unsigned long long>(run))) : 0;
  }
  __syncthreads();
  int slot = block_base + wave_total[w_id] + incl - mine;
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
  if (threadIdx.x == 0 && wsum) atomicAdd(weight_sum, wsum);
}

/// Path costs, first step: every vertex's own tree edge (p, v) as the first alive (p, v) of p's
/// out-row (the base run, then the tail: MOSP's row order), its first `columns` weights; the
/// source starts at 0 and an unreachable vertex at INF, both with no ancestor. A tree edge that
/// is not an edge of the graph is reported (the lowest such vertex).
__global__ void k_tree_edge_costs(const int2* meta, const int32_t* base, const int32_t* nbr, const int32_t* w,
                                  int64_t slot_cap, int columns, const int32_t* parents, int32_t n, int32_t source,
                                  int64_t* acc, int32_t* jump, int* missing) {
  const int32_t v = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (v >= n) return;
  const int32_t p = v == source ? -1 : parents[v];
  int64_t* a = acc + static_cast<int64_t>(v) * columns;
  if (p < 0) {
    const int64_t start = v == source ? 0 : sssp_infinity;
    for (int c = 0; c < columns; ++c) a[c] = start;
    jump[v] = -1;
    return;
  }
  const int2 mt = meta[p];
  const int32_t* row = nbr + mt.x;
  const int bl = base[p];
  int hit = -1;
  for (int r = 0; r < 2 && hit < 0; ++r) {
    const int lo = r == 0 ? 0 : bl, hi = r == 0 ? bl : mt.y;
    for (int x = row_lower_bound(row, lo, hi, v); x < hi && (row[x] & 0x7fffffff) == v; ++x) {
      if (row[x] >= 0) {
        hit = x;
        break;
      }
    }
  }
  if (hit < 0) {
    atomicMin(missing, v);
    for (int c = 0; c < columns; ++c) a[c] = sssp_infinity;
    jump[v] = -1;
    return;
  }
  for (int c = 0; c < columns; ++c) a[c] = w[c * slot_cap + mt.x + hit];
  jump[v] = p;
}

/// Path doubling: each vertex adds its ancestor's sum and jumps to the ancestor's ancestor, so
/// after ceil(log2 n) + 1 rounds every sum runs back to the source.
__global__ void k_path_double(const int64_t* acc_in, const int32_t* jump_in, int64_t* acc_out, int32_t* jump_out,
                              int32_t n, int columns) {
  const int32_t v = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x);
  if (v >= n) return;
  const int32_t j = jump_in[v];
  const int64_t* a = acc_in + static_cast<int64_t>(v) * columns;
  int64_t* o = acc_out + static_cast<int64_t>(v) * columns;
  if (j < 0) {
    for (int c = 0; c < columns; ++c) o[c] = a[c];
    jump_out[v] = -1;
    return;
  }
  const int64_t* b = acc_in + static_cast<int64_t>(j) * columns;
  for (int c = 0; c < columns; ++c) o[c] = a[c] + b[c];
  jump_out[v] = jump_in[j];
}

unsigned int grid(int64_t work, int block) { return static_cast<unsigned int>(std::max<int64_t>(1, (work + block - 1) / block)); }

void check_status(int status) {
  if (status != DYNG_HIP_OK) throw status_error(status, dyng_hip_last_error());
}

}  // namespace

dyng_hip_graph* mosp_build_combined(int device, const int32_t* const* parents, int32_t num_objectives, int64_t n,
                                    int32_t source, int32_t base, const int32_t* terms, hipStream_t s,
                                    int64_t* edges, int64_t* weight_sum) {
  dbuf<const int32_t*> dparents;
  dbuf<int32_t> dterms;
  dbuf<int32_t> esrc, edst, ew;
  dbuf<unsigned long long> totals;
  dparents.reserve(num_objectives);
  dterms.reserve(num_objectives);
  const int64_t cap = std::max<int64_t>(1, n * num_objectives);
  esrc.reserve(cap);
  edst.reserve(cap);
  ew.reserve(cap);
  totals.reserve(2);
  // small synchronous copies from the caller's host arrays (done before anything is queued)
  DYNG_HIP_TRY(hipMemcpy(dparents.p, parents, num_objectives * sizeof(const int32_t*), hipMemcpyHostToDevice));
  DYNG_HIP_TRY(hipMemcpy(dterms.p, terms, num_objectives * sizeof(int32_t), hipMemcpyHostToDevice));
  DYNG_HIP_TRY(hipMemsetAsync(totals.p, 0, 2 * sizeof(unsigned long long), s));
  combine_params c{dparents.p, dterms.p, num_objectives, static_cast<int32_t>(n), source, base};
  hipLaunchKernelGGL(k_combine, dim3(grid(n, kBlock)), dim3(kBlock), 0, s, c, esrc.p, edst.p, ew.p, totals.p,
                     totals.p + 1);
  DYNG_HIP_TRY(hipGetLastError());
  unsigned long long host[2] = {0, 0};
  DYNG_HIP_TRY(hipMemcpyAsync(host, totals.p, sizeof(host), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  *edges = static_cast<int64_t>(host[0]);
  *weight_sum = static_cast<int64_t>(host[1]);
  // a simple directed graph with sorted rows; the combined edges are distinct pairs already
  dyng_hip_graph_props props{1, 1, 0, 0, 0, 0, 1, 1, 0.25};
  dyng_hip_graph* cg = nullptr;
  check_status(dyng_hip_graph_create_from_edges(device, &props, n, *edges, 1, esrc.p, edst.p, ew.p, 1, s, &cg));
  return cg;
}

void mosp_path_costs(const dyng_hip_graph* g, const int32_t* parents, int64_t n, int32_t source, int32_t columns,
                     int64_t* costs, hipStream_t s) {
  if (n == 0 || columns == 0) return;
  dbuf<int64_t> acc;
  dbuf<int32_t> jump_a, jump_b;
  dbuf<int> missing;
  acc.reserve(static_cast<size_t>(n) * columns);
  jump_a.reserve(n);
  jump_b.reserve(n);
  missing.reserve(1);
  const int first_missing = INT32_MAX;
  DYNG_HIP_TRY(hipMemcpy(missing.p, &first_missing, sizeof(int), hipMemcpyHostToDevice));
  // the first round writes the caller's array; the rounds then alternate between it and `acc`
  int64_t* bufs[2] = {costs, acc.p};
  int32_t* jumps[2] = {jump_a.p, jump_b.p};
  hipLaunchKernelGGL(k_tree_edge_costs, dim3(grid(n, 256)), dim3(256), 0, s, g->out.meta.p, g->out.base.p, g->out.nbr.p,
                     g->out.w.p, g->out.slot_cap, columns, parents, static_cast<int32_t>(n), source, bufs[0], jumps[0],
                     missing.p);
  int rounds = 1;
  while ((int64_t{1} << rounds) < n) ++rounds;
  rounds += 1;
  int cur = 0;
  for (int r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(k_path_double, dim3(grid(n, 256)), dim3(256), 0, s, bufs[cur], jumps[cur], bufs[cur ^ 1],
                       jumps[cur ^ 1], static_cast<int32_t>(n), columns);
    cur ^= 1;
  }
  DYNG_HIP_TRY(hipGetLastError());
  if (cur != 0) {
    DYNG_HIP_TRY(hipMemcpyAsync(costs, acc.p, static_cast<size_t>(n) * columns * sizeof(int64_t),
                                hipMemcpyDeviceToDevice, s));
  }
  int host_missing = INT32_MAX;
  DYNG_HIP_TRY(hipMemcpyAsync(&host_missing, missing.p, sizeof(int), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (host_missing != INT32_MAX) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "mosp: the MOSP tree edge (" + std::to_string(host_missing) +
                                        "'s parent, " + std::to_string(host_missing) +
                                        ") is not an edge of the graph (do the trees belong to this graph?)");
  }
}

}  // namespace dyng_hip
