// The device graph of the HIP sssp path: out- and in-rows in slack storage, built on the device
// and updated in place by each batch.
//
// Layout (HBM): per vertex an int2 {offset, length}, a capacity, the length of the row's base run
// and the count of dead entries in it; one slot arena per direction holding the neighbour ids,
// then one int32 weight column per objective (column k at w + k * slot_cap). A row is two sorted
// runs: the base [offset, offset + base) and the tail [offset + base, offset + length), each
// sorted by neighbour id with parallel edges in the order they were added; among equal
// neighbours the base entry is the older one. A deletion marks an entry dead (the sign bit of its
// neighbour id) instead of moving the row. So the first alive (u,v) of row u in MOSP's append
// order is the first alive entry of v's run in the base, else in the tail. A batch rewrites only
// a row's tail while the tail stays short and the base has few dead entries (tail_limit,
// dead_limit); otherwise the whole row is merged into one new base run. A row's capacity is
// len + ceil(len * headroom) + 2; a row that outgrows it moves to the free tail of the arena;
// compaction rebuilds the arena when the tail runs short.
//
// A batch (MOSP applyChangeBatch under upsert_last_wins, cpp/src/graph/apply_host.cpp:57-401) is
// applied without leaving the device:
//   1. Step 0 (apply_host.cpp:80-147): self-loops, vertex growth, undirected expansion; one key
//      (u << b | v) per effective operation, in application order (deletions first).
//   2. A stable radix sort by key groups the operations of each pair (u,v) in batch order.
//   3. One thread per pair replays its operations on the runs of v in row u (binary searches in
//      both runs): deletions mark the first alive match dead, insertions overwrite the first
//      alive match (upsert) or add a pending entry; then the weight-increase flags of :348-394
//      (final first (u,v) against the original first (u,v)). Each operation that changes an
//      in-edge emits an in-operation for the in-rows.
//   4. A planning pass sorts the touched rows into three merge modes (a short tail merged by one
//      wave in registers, a longer tail by one workgroup, or the whole row into a new base) and
//      sums the staging and relocation slots the full merges need, so the host can grow or
//      compact the arena first; then one launch per mode merges the alive entries with the
//      pending ones (positions from prefix sums and binary searches), in place, or at the arena's
//      tail when the row outgrew its capacity.
//   5. Steps 2-4 again for the in-rows with the in-operations (a deletion or a weight change
//      matches the first alive in-entry of u with the old weights).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace dyng_hip {

namespace {

constexpr int kBlock = 256;
constexpr int32_t kDead = static_cast<int32_t>(0x80000000u);
constexpr u64 kNoKey = ~0ULL;

unsigned int blocks_for(int64_t work, int block = kBlock) {
  int64_t b = (work + block - 1) / block;
  if (b < 1) b = 1;
  if (b > 65535LL * 16) b = 65535LL * 16;
  return static_cast<unsigned int>(b);
}

int bits_for(int64_t n_after) {
  int b = 1;
  while ((int64_t{1} << b) <= n_after) ++b;  // n_after < 2^b, so ids < 2^b - 1 < the all-ones key
  return b;
}

DYNG_DEV int64_t lower_key(const u64* keys, int64_t lo, int64_t hi, u64 key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// ---------------------------------------------------------------------------------------------
// rocPRIM wrappers
// ---------------------------------------------------------------------------------------------

void sort_pairs(dbuf<unsigned char>& tmp, const u64* kin, u64* kout, const uint32_t* vin,
                uint32_t* vout, size_t n, int end_bit, hipStream_t s) {
  if (n == 0) return;
  size_t bytes = 0;
  DYNG_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, kin, kout, vin, vout, n, 0, end_bit, s));
  tmp.reserve(bytes);
  DYNG_HIP_TRY(rocprim::radix_sort_pairs(static_cast<void*>(tmp.p), bytes, kin, kout, vin, vout, n, 0,
                                         end_bit, s));
}

template <typename T>
void exclusive_scan(dbuf<unsigned char>& tmp, const T* in, T* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  size_t bytes = 0;
  DYNG_HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, T{0}, n, rocprim::plus<T>(), s));
  tmp.reserve(bytes);
  DYNG_HIP_TRY(rocprim::exclusive_scan(static_cast<void*>(tmp.p), bytes, in, out, T{0}, n,
                                       rocprim::plus<T>(), s));
}

/// The exclusive prefix sums of four int32 arrays of n entries (the pending and killed flags of
/// both directions' operations), in two launches: the sums of tiles of kScan4Tile entries, then
/// each tile scanned from the sum of the tiles before it (each workgroup adds those up itself).
/// No chain of look-backs between workgroups: on a few hundred thousand entries a decoupled
/// look-back scan is a serial chain of device-scope round trips (rocPRIM's took 16 us here, its
/// state initialisation included).
constexpr int kScan4Block = 1024;
constexpr int kScan4Items = 1;  ///< entries per thread: tiles of 1024 (four per thread measured 5 us slower)
constexpr int kScan4Tile = kScan4Block * kScan4Items;

struct scan4_args {
  const int32_t* in[4];
  int32_t* out[4];
  int64_t n;
  int4* tile_sums;
};

/// Sum over the workgroup of one int4 per thread; thread 0 gets it.
DYNG_DEV int4 block_sum4(int4 v, int4* part) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
    v.z += __shfl_xor(v.z, off);
    v.w += __shfl_xor(v.w, off);
  }
  if (lane() == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  int4 t = make_int4(0, 0, 0, 0);
  if (threadIdx.x == 0) {
    for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) {
      t.x += part[w].x;
      t.y += part[w].y;
      t.z += part[w].z;
      t.w += part[w].w;
    }
  }
  return t;
}

__global__ void __launch_bounds__(kScan4Block) k_scan4_tiles(scan4_args a) {
  __shared__ int4 part[kScan4Block / 64];
  const int64_t t0 = static_cast<int64_t>(blockIdx.x) * kScan4Tile;
  int4 v = make_int4(0, 0, 0, 0);
  for (int j = 0; j < kScan4Items; ++j) {
    const int64_t i = t0 + static_cast<int64_t>(j) * kScan4Block + threadIdx.x;
    if (i < a.n) {
      v.x += a.in[0][i];
      v.y += a.in[1][i];
      v.z += a.in[2][i];
      v.w += a.in[3][i];
    }
  }
  const int4 t = block_sum4(v, part);
  if (threadIdx.x == 0) a.tile_sums[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kScan4Block) k_scan4_apply(scan4_args a) {
  __shared__ int4 part[kScan4Block / 64];
  __shared__ int4 wave_sum[kScan4Block / 64];
  __shared__ int4 base;
  // the sum of the tiles before this one
  int4 v = make_int4(0, 0, 0, 0);
  for (int b = threadIdx.x; b < static_cast<int>(blockIdx.x); b += blockDim.x) {
    const int4 x = a.tile_sums[b];
    v.x += x.x;
    v.y += x.y;
    v.z += x.z;
    v.w += x.w;
  }
  const int4 before = block_sum4(v, part);
  if (threadIdx.x == 0) base = before;
  // this thread's kScan4Items consecutive entries, their sums, then the workgroup's exclusive scan
  const int64_t first = static_cast<int64_t>(blockIdx.x) * kScan4Tile + static_cast<int64_t>(threadIdx.x) * kScan4Items;
  int32_t x[4][kScan4Items];
  int4 own = make_int4(0, 0, 0, 0);
#pragma unroll
  for (int j = 0; j < kScan4Items; ++j) {
    const int64_t i = first + j;
    const bool in = i < a.n;
    x[0][j] = in ? a.in[0][i] : 0;
    x[1][j] = in ? a.in[1][i] : 0;
    x[2][j] = in ? a.in[2][i] : 0;
    x[3][j] = in ? a.in[3][i] : 0;
    own.x += x[0][j];
    own.y += x[1][j];
    own.z += x[2][j];
    own.w += x[3][j];
  }
  int4 inc = own;  // inclusive scan over the wave
  for (int off = 1; off < 64; off <<= 1) {
    const int4 y = make_int4(__shfl_up(inc.x, off), __shfl_up(inc.y, off), __shfl_up(inc.z, off), __shfl_up(inc.w, off));
    if (lane() >= off) {
      inc.x += y.x;
      inc.y += y.y;
      inc.z += y.z;
      inc.w += y.w;
    }
  }
  const int wv = static_cast<int>(threadIdx.x / 64);
  if (lane() == 63) wave_sum[wv] = inc;
  __syncthreads();  // wave sums and base are in LDS
  int4 run = base;
  for (int w = 0; w < wv; ++w) {
    run.x += wave_sum[w].x;
    run.y += wave_sum[w].y;
    run.z += wave_sum[w].z;
    run.w += wave_sum[w].w;
  }
  run.x += inc.x - own.x;
  run.y += inc.y - own.y;
  run.z += inc.z - own.z;
  run.w += inc.w - own.w;
#pragma unroll
  for (int j = 0; j < kScan4Items; ++j) {
    const int64_t i = first + j;
    if (i < a.n) {
      a.out[0][i] = run.x;
      a.out[1][i] = run.y;
      a.out[2][i] = run.z;
      a.out[3][i] = run.w;
    }
    run.x += x[0][j];
    run.y += x[1][j];
    run.z += x[2][j];
    run.w += x[3][j];
  }
}

void exclusive_scan4(dbuf<unsigned char>& tmp, const int32_t* a, const int32_t* b, const int32_t* c, const int32_t* d,
                     int32_t* oa, int32_t* ob, int32_t* oc, int32_t* od, size_t n, hipStream_t s) {
  if (n == 0) return;
  const int64_t tiles = (static_cast<int64_t>(n) + kScan4Tile - 1) / kScan4Tile;
  tmp.reserve(static_cast<size_t>(tiles) * sizeof(int4));
  scan4_args args{{a, b, c, d}, {oa, ob, oc, od}, static_cast<int64_t>(n), reinterpret_cast<int4*>(tmp.p)};
  hipLaunchKernelGGL(k_scan4_tiles, dim3(static_cast<unsigned>(tiles)), dim3(kScan4Block), 0, s, args);
  hipLaunchKernelGGL(k_scan4_apply, dim3(static_cast<unsigned>(tiles)), dim3(kScan4Block), 0, s, args);
  DYNG_HIP_TRY(hipGetLastError());
}

template <typename T>
__global__ void k_fill(T* p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < n; i += int64_t{gridDim.x} * blockDim.x) p[i] = v;
}

template <typename T>
void fill(T* p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<T>, dim3(blocks_for(n)), dim3(kBlock), 0, s, p, n, v);
  DYNG_HIP_TRY(hipGetLastError());
}

// ---------------------------------------------------------------------------------------------
// Construction
// ---------------------------------------------------------------------------------------------

/// One key per (expanded) input edge: (u << b | v), or kNoKey for a dropped self-loop or the
/// missing reverse arc of a directed edge. val = slot (edge index * stride + reversed).
__global__ void k_edge_keys(const int32_t* src, const int32_t* dst, int64_t count, int stride,
                            int directed, int drop_self_loops, int vbits, u64* keys, uint32_t* vals) {
  const int64_t total = count * stride;
  for (int64_t e = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; e < total; e += int64_t{gridDim.x} * blockDim.x) {
    const int64_t j = e / stride;
    const int r = static_cast<int>(e % stride);
    const int32_t u0 = src[j], v0 = dst[j];
    bool valid = true;
    if (r == 1 && (directed || u0 == v0)) valid = false;
    if (u0 == v0 && drop_self_loops) valid = false;
    const int32_t u = r ? v0 : u0, v = r ? u0 : v0;
    keys[e] = valid ? ((static_cast<u64>(u) << vbits) | static_cast<u64>(v)) : kNoKey;
    vals[e] = static_cast<uint32_t>(e);
  }
}

/// keep[i]: sorted edge i survives (a simple graph keeps the last of a run of equal pairs).
__global__ void k_keep(const u64* keys, int64_t N, int allow_parallel, int32_t* keep) {
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i <= N; i += int64_t{gridDim.x} * blockDim.x) {
    if (i == N) { keep[i] = 0; continue; }
    const u64 k = keys[i];
    keep[i] = (k != kNoKey && (allow_parallel || i + 1 == N || keys[i + 1] != k)) ? 1 : 0;
  }
}

/// Degree and capacity of every row from the sorted keys and the keep prefix.
__global__ void k_row_degrees(const u64* keys, int64_t N, const int32_t* keep_prefix, int64_t n,
                              int vbits, double headroom, int2* meta, int32_t* cap, int64_t* cap64) {
  for (int64_t u = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; u < n; u += int64_t{gridDim.x} * blockDim.x) {
    const int64_t a = lower_key(keys, 0, N, static_cast<u64>(u) << vbits);
    const int64_t b = lower_key(keys, a, N, static_cast<u64>(u + 1) << vbits);
    const int deg = keep_prefix[b] - keep_prefix[a];
    const int c = deg + static_cast<int>(ceil(deg * headroom)) + 2;
    meta[u] = make_int2(0, deg);
    cap[u] = c;
    cap64[u] = c;
  }
}

__global__ void k_set_offsets(const int64_t* off64, int64_t n, int2* meta) {
  for (int64_t u = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; u < n; u += int64_t{gridDim.x} * blockDim.x) {
    meta[u].x = static_cast<int32_t>(off64[u]);
  }
}

/// Scatter the kept sorted edges into their rows. Weights come from the input list
/// (w_in[edge * K + k], edge = slot / stride) or, when `from_slots`, from an arena column
/// (w_in[k * in_slot_cap + slot]).
__global__ void k_fill_rows(const u64* keys, const uint32_t* vals, int64_t N, const int32_t* keep,
                            const int32_t* keep_prefix, int vbits, int stride, const int2* meta,
                            int32_t* nbr, int32_t* w, int64_t slot_cap, int K, const int32_t* w_in,
                            int from_slots, int64_t in_slot_cap) {
  const u64 mask = (1ULL << vbits) - 1;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < N; i += int64_t{gridDim.x} * blockDim.x) {
    if (!keep[i]) continue;
    const u64 key = keys[i];
    const int64_t u = static_cast<int64_t>(key >> vbits);
    const int32_t v = static_cast<int32_t>(key & mask);
    const int64_t a = lower_key(keys, 0, i + 1, static_cast<u64>(u) << vbits);
    const int64_t slot = meta[u].x + (keep_prefix[i] - keep_prefix[a]);
    nbr[slot] = v;
    const uint32_t val = vals[i];
    for (int k = 0; k < K; ++k) {
      w[k * slot_cap + slot] = from_slots ? w_in[k * in_slot_cap + val] : w_in[(static_cast<int64_t>(val) / stride) * K + k];
    }
  }
}

/// Keys of the in-rows from the out-rows: (v << b | u), val = out slot.
__global__ void k_reverse_keys(const int2* meta, const int32_t* nbr, int64_t n, int vbits, u64* keys,
                               uint32_t* vals, const int64_t* row_pos) {
  // one wave per row
  const int64_t wave = (blockIdx.x * int64_t{blockDim.x} + threadIdx.x) / 64;
  const int64_t waves = (int64_t{gridDim.x} * blockDim.x) / 64;
  for (int64_t u = wave; u < n; u += waves) {
    const int2 mt = meta[u];
    const int64_t base = row_pos[u];
    for (int j = lane(); j < mt.y; j += 64) {
      const int32_t v = nbr[mt.x + j];
      keys[base + j] = (static_cast<u64>(v) << vbits) | static_cast<u64>(u);
      vals[base + j] = static_cast<uint32_t>(mt.x + j);
    }
  }
}

__global__ void k_len64(const int2* meta, int64_t n, int64_t* out) {
  for (int64_t u = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; u < n; u += int64_t{gridDim.x} * blockDim.x) out[u] = meta[u].y;
}

/// Column summary of every objective: pass 0 max and sum, pass 1 the count of maximal weights.
/// Each wave takes 64 consecutive rows at a time: a short row is scanned by its own lane, a long
/// one by the whole wave (a wave per row would leave most lanes idle on sparse rows and chain one
/// row's latency after another).
constexpr int kShortRow = 16;

__global__ void k_column_summary(const int2* meta, const int32_t* nbr, const int32_t* w, int64_t n, int64_t slot_cap,
                                 int K, int pass, apply_control* ctl) {
  const int64_t wave = (blockIdx.x * int64_t{blockDim.x} + threadIdx.x) / 64;
  const int64_t waves = (int64_t{gridDim.x} * blockDim.x) / 64;
  for (int k = 0; k < K; ++k) {
    long long mx = 0, sum = 0, cnt = 0, bad = 0;
    const long long target = ctl->col_max[k];
    const int32_t* col = w + k * slot_cap;
    auto visit = [&](int64_t slot) {
      if (nbr[slot] < 0) return;  // dead
      const long long x = col[slot];
      if (pass == 0) {
        mx = x > mx ? x : mx;
        sum += x;
        bad += x < 1 ? 1 : 0;
      } else {
        cnt += x == target ? 1 : 0;
      }
    };
    for (int64_t r0 = wave * 64; r0 < n; r0 += waves * 64) {
      const int64_t u = r0 + lane();
      const int2 mt = u < n ? meta[u] : make_int2(0, 0);
      const bool is_long = mt.y > kShortRow;
      if (!is_long) {
        for (int j = 0; j < mt.y; ++j) visit(mt.x + int64_t{j});
      }
      u64 longs = __ballot(is_long);
      while (longs) {
        const int l = __ffsll(static_cast<unsigned long long>(longs)) - 1;
        longs &= longs - 1;
        const int off = __shfl(mt.x, l), len = __shfl(mt.y, l);
        for (int j = lane(); j < len; j += 64) visit(off + int64_t{j});
      }
    }
    if (pass == 0) {
      mx = static_cast<long long>(wave_min(~static_cast<u64>(mx)));
      mx = static_cast<long long>(~static_cast<u64>(mx));
      sum = static_cast<long long>(wave_sum(static_cast<u64>(sum)));
      bad = static_cast<long long>(wave_sum(static_cast<u64>(bad)));
      if (lane() == 0) {
        atomicMax(reinterpret_cast<unsigned long long*>(&ctl->col_max[k]), static_cast<unsigned long long>(mx));
        atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->col_dsum[k]), static_cast<unsigned long long>(sum));
        if (bad) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->col_bad[k]), static_cast<unsigned long long>(bad));
      }
    } else {
      cnt = static_cast<long long>(wave_sum(static_cast<u64>(cnt)));
      if (lane() == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&ctl->col_dcount[k]), static_cast<unsigned long long>(cnt));
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Apply
// ---------------------------------------------------------------------------------------------

/// Batch scan (before anything changes): id range, self-loops, weight range per objective. The
/// insertions and deletions are read in one loop (their loads in flight together); each
/// workgroup reduces its nine values in one pass (wave shuffles, one barrier) and adds them to
/// the control block with one atomic per value: the launch is 64 workgroups of 1024 threads, so
/// few workgroups queue on those nine addresses, which share a cache line.
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_batch_scan(const int32_t* ins_src, const int32_t* ins_dst,
                                                            const int32_t* ins_w, int64_t num_ins,
                                                            const int32_t* del_src, const int32_t* del_dst,
                                                            int64_t num_del, int K, apply_control* ctl) {
  long long mn = INT32_MAX, mx = -1, sl_i = 0, sl_d = 0, wmin = INT32_MAX;
  long long lmax0 = 0, lmax1 = 0, lmax2 = 0, lmax3 = 0;  // named: a private array would sit in scratch
  const int64_t stride = int64_t{gridDim.x} * blockDim.x;
  const int64_t total = num_ins > num_del ? num_ins : num_del;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < total; i += stride) {
    const bool ins = i < num_ins, del = i < num_del;
    const int32_t u = ins ? ins_src[i] : 0, v = ins ? ins_dst[i] : 0;
    const int32_t du = del ? del_src[i] : 0, dv = del ? del_dst[i] : 0;
    if (ins) {
      mn = min(mn, static_cast<long long>(min(u, v)));
      mx = max(mx, static_cast<long long>(max(u, v)));
      sl_i += u == v ? 1 : 0;
      for (int k = 0; k < K; ++k) {
        const int32_t x = ins_w[i * K + k];
        wmin = min(wmin, static_cast<long long>(x));
        if (k == 0) lmax0 = max(lmax0, static_cast<long long>(x));
        else if (k == 1) lmax1 = max(lmax1, static_cast<long long>(x));
        else if (k == 2) lmax2 = max(lmax2, static_cast<long long>(x));
        else if (k == 3) lmax3 = max(lmax3, static_cast<long long>(x));
        else atomicMax(&ctl->batch_wmax[k], x);
        if (x < 1) atomicMin(&ctl->batch_wmin[k], x);  // rare: the objective checks are the host's
      }
    }
    if (del) {
      mn = min(mn, static_cast<long long>(min(du, dv)));
      sl_d += du == dv ? 1 : 0;
    }
  }
  constexpr int kVals = 9;
  long long val[kVals] = {mn, mx, sl_i, sl_d, wmin, lmax0, lmax1, lmax2, lmax3};
  auto combine = [](int q, long long a, long long b) -> long long {
    if (q == 0 || q == 4) return b < a ? b : a;  // minima
    if (q == 2 || q == 3) return a + b;          // counts
    return b > a ? b : a;                        // maxima
  };
#pragma unroll
  for (int q = 0; q < kVals; ++q) {
    for (int off = 32; off > 0; off >>= 1) val[q] = combine(q, val[q], static_cast<long long>(__shfl_xor(val[q], off)));
  }
  __shared__ long long part[kScanBlock / 64][kVals];
  if (lane() == 0) {
#pragma unroll
    for (int q = 0; q < kVals; ++q) part[threadIdx.x / 64][q] = val[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int waves = static_cast<int>(blockDim.x / 64);
#pragma unroll
    for (int q = 0; q < kVals; ++q) {
      long long t = part[0][q];
      for (int w = 1; w < waves; ++w) t = combine(q, t, part[w][q]);
      val[q] = t;
    }
    if (val[0] != INT32_MAX) atomicMin(&ctl->min_id, static_cast<int>(val[0]));
    if (val[1] >= 0) atomicMax(&ctl->max_id, static_cast<int>(val[1]));
    if (val[2]) atomicAdd(&ctl->self_loops_ins, static_cast<int>(val[2]));
    if (val[3]) atomicAdd(&ctl->self_loops_del, static_cast<int>(val[3]));
    if (val[4] != INT32_MAX) atomicMin(&ctl->min_weight, static_cast<int>(val[4]));
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k < K && val[5 + k] > 0) atomicMax(&ctl->batch_wmax[k], static_cast<int>(val[5 + k]));
    }
  }
}

struct op_layout {
  int64_t num_first, num_second;  ///< batch operations of the first / second kind
  int first_is_del;               ///< deletions_first
  int stride;                     ///< 1 directed, 2 undirected (reverse arc slot)
  DYNG_DEV bool is_del(int64_t e) const {
    return (e < num_first * stride) ? first_is_del != 0 : first_is_del == 0;
  }
  DYNG_DEV int64_t local(int64_t e) const { return e < num_first * stride ? e : e - num_first * stride; }
  DYNG_DEV int64_t del_base() const { return first_is_del ? 0 : num_first * stride; }
  DYNG_DEV int64_t ins_base() const { return first_is_del ? num_first * stride : 0; }
};

/// Ranks of offending operations in the error modes, in the order the reference's host apply
/// meets them (apply_host.cpp:80-252): first the deletions naming a vertex past the vertex count
/// (batch order), then row by row (ascending u), within a row the first phase's operations (the
/// deletions when deletions_first) before the second phase's, each in batch order. The lowest
/// rank is the error reported. Layout: bit 63 row phase, bits 62-32 the row, bit 31 the second
/// phase, bits 30-0 the operation's batch index.
constexpr unsigned long long kNoError = ~0ULL;
DYNG_DEV unsigned long long error_rank_prescan(int64_t j) { return static_cast<unsigned long long>(j); }
DYNG_DEV unsigned long long error_rank_row(int32_t u, int second_phase, int64_t j) {
  return (1ULL << 63) | (static_cast<unsigned long long>(u) << 32) | (static_cast<unsigned long long>(second_phase) << 31) |
         static_cast<unsigned long long>(j);
}

__global__ void k_make_ops(op_layout L, const int32_t* ins_src, const int32_t* ins_dst,
                           const int32_t* del_src, const int32_t* del_dst, int directed,
                           int drop_self_loops, int64_t n_after, int vbits, u64* keys, uint32_t* vals,
                           int32_t* e_del_src, int32_t* e_del_dst, int32_t* e_ins_src,
                           int32_t* e_ins_dst, uint8_t* e_inc, int32_t* e_fw, int K, int missing_delete_error,
                           apply_control* ctl) {
  const int64_t total = (L.num_first + L.num_second) * L.stride;
  unsigned long long dropped = 0, ignored = 0;
  for (int64_t e = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; e < total; e += int64_t{gridDim.x} * blockDim.x) {
    const bool del = L.is_del(e);
    const int64_t loc = L.local(e);
    const int64_t j = loc / L.stride;
    const int r = static_cast<int>(loc % L.stride);
    const int32_t u0 = del ? del_src[j] : ins_src[j];
    const int32_t v0 = del ? del_dst[j] : ins_dst[j];
    bool valid = true;
    if (r == 1 && (directed || u0 == v0)) valid = false;
    if (valid && u0 == v0 && drop_self_loops) {
      valid = false;
      dropped += 1;
    }
    const int32_t u = r ? v0 : u0, v = r ? u0 : v0;
    if (valid && del && (u >= n_after || v >= n_after)) {
      valid = false;
      ignored += 1;
      // no such vertex, so no such edge: in the error mode, the batch's first such deletion
      // (apply_host.cpp:116-138 finds these before any row)
      if (missing_delete_error) atomicMin(&ctl->first_error, error_rank_prescan(j));
    }
    // Both directions' keys go through one sort: the out-row key (u, v) at e and the in-row key
    // (v, u) at total + e, tagged with bit 2 * vbits so that every out-key sorts before every
    // in-key (what an operation does to the in-row is decided later by the out-row pass). A
    // dropped operation takes the all-ones key of its half (no vertex id is all ones); k_untag
    // restores kNoKey and removes the tag after the sort.
    const u64 low = (u64{1} << (2 * vbits)) - 1;
    const u64 tag = u64{1} << (2 * vbits);
    keys[e] = valid ? ((static_cast<u64>(u) << vbits) | static_cast<u64>(v)) : low;
    vals[e] = static_cast<uint32_t>(e);
    keys[total + e] = tag | (valid ? ((static_cast<u64>(v) << vbits) | static_cast<u64>(u)) : low);
    vals[total + e] = static_cast<uint32_t>(e);
    if (del) {
      e_del_src[loc] = valid ? u : -1;
      e_del_dst[loc] = valid ? v : -1;
    } else {
      e_ins_src[loc] = valid ? u : -1;
      e_ins_dst[loc] = valid ? v : -1;
      for (int k = 0; k < K; ++k) e_inc[loc * K + k] = 0;
      if (e_fw != nullptr) {
        for (int k = 0; k < K; ++k) e_fw[loc * K + k] = -1;
      }
    }
  }
  dropped = wave_sum(dropped);
  ignored = wave_sum(ignored);
  if (lane() == 0) {
    if (dropped) atomicAdd(&ctl->dropped_self_loops, dropped);
    if (ignored) atomicAdd(&ctl->ignored_deletions, ignored);
  }
}

/// After the joint sort: the keys without their direction tag, kNoKey for dropped operations.
__global__ void k_untag(u64* keys, int64_t count, int vbits) {
  const u64 low = (u64{1} << (2 * vbits)) - 1;
  for (int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; i < count; i += int64_t{gridDim.x} * blockDim.x) {
    const u64 k = keys[i] & low;
    keys[i] = k == low ? kNoKey : k;
  }
}

/// Rows up to this many entries are merged by one wave, longer ones by a workgroup.
constexpr int kMergeSmall = 512;

struct out_pair_params {
  op_layout L;
  const u64* keys;
  const uint32_t* vals;
  int64_t N;
  int vbits;
  int K;
  int64_t n_before;  ///< rows that existed before the batch (classification)
  int upsert;        ///< on_existing_insert: 1 upsert, 0 ignore
  const int32_t* ins_w;  ///< batch insertion weights (j * K + k)
  int2* meta;
  const int32_t* base;
  int32_t* dead;
  int32_t* nbr;
  int32_t* w;
  int64_t slot_cap;
  int32_t* pend_alive;
  int32_t* kill;    ///< 1: the operation killed a tail entry
  int32_t* pend_w;  ///< k * N + position
  uint8_t* e_inc;
  /// simple graphs: per insertion (batch slot, K weights) the pair's weights after the batch, -1
  /// if the pair is gone (null with parallel edges, where several entries may hold the pair)
  int32_t* e_fw;
  // what each operation (by batch slot e) does to the in-rows: kInNone, or the type and weights
  int32_t* iop_type;
  int32_t* iop_wold;
  int32_t* iop_wnew;
  int32_t* mark;
  int mark_gen;
  int32_t* touched;
  int32_t* row_a;  ///< per vertex: first operation of its row in the sorted keys
  int32_t* row_b;  ///< per vertex: one past its last operation
  apply_control* ctl;
};

enum { kInDelete = 0, kInModify = 1, kInInsert = 2, kInNone = 3 };

/// The entries of neighbour v in a row's two sorted runs: [lb, ub) in the base, [tl, tu) in the
/// tail (row-relative indices).
struct run_pair {
  int bl, lb, ub, tl, tu;
};

DYNG_DEV run_pair find_runs(const int32_t* row, int base_len, int len, int32_t v) {
  run_pair r;
  r.bl = base_len;
  r.lb = row_lower_bound(row, 0, base_len, v);
  r.ub = row_upper_bound_from(row, r.lb, base_len, v);
  r.tl = row_lower_bound(row, base_len, len, v);
  r.tu = row_upper_bound_from(row, r.tl, len, v);
  return r;
}

/// The error modes, checked before anything changes (read only): one thread per (u, v) segment of
/// the sorted out-row keys replays the segment's operations in batch order on counts of the alive
/// (u, v) entries (old ones in the row, pending ones of the batch). A deletion that finds none is a
/// missing edge, an insertion that finds one an existing edge (apply_host.cpp:117-120, 206-226);
/// the lowest rank of an offending operation lands in ctl->first_error.
__global__ void __launch_bounds__(kBlock) k_check_pairs(op_layout L, const u64* keys, const uint32_t* vals,
                                                         int64_t N, int vbits, int64_t n_before, const int2* meta,
                                                         const int32_t* base, const int32_t* nbr,
                                                         int missing_delete_error, int existing_insert_error,
                                                         apply_control* ctl) {
  const int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x;
  if (i >= N) return;
  const u64 key = keys[i];
  if (key == kNoKey || (i > 0 && keys[i - 1] == key)) return;
  const u64 mask = (1ULL << vbits) - 1;
  const int32_t u = static_cast<int32_t>(key >> vbits);
  const int32_t v = static_cast<int32_t>(key & mask);
  int alive = 0;
  if (u < n_before) {
    const int2 mt = meta[u];
    const int32_t* row = nbr + mt.x;
    const run_pair rr = find_runs(row, base[u], mt.y, v);
    for (int x = rr.lb; x < rr.ub; ++x) alive += row[x] >= 0 ? 1 : 0;
    for (int x = rr.tl; x < rr.tu; ++x) alive += row[x] >= 0 ? 1 : 0;
  }
  unsigned long long first = kNoError;
  for (int64_t q = i; q < N && keys[q] == key; ++q) {
    const int64_t e = vals[q];
    const int second = e >= L.num_first * L.stride ? 1 : 0;
    const int64_t j = L.local(e) / L.stride;
    if (L.is_del(e)) {
      if (alive > 0) {
        --alive;
      } else if (missing_delete_error) {
        first = min(first, error_rank_row(u, second, j));
      }
    } else if (alive > 0) {
      if (existing_insert_error) first = min(first, error_rank_row(u, second, j));
    } else {
      ++alive;  // appended
    }
  }
  if (first != kNoError) atomicMin(&ctl->first_error, first);
}

/// Record where row r's operations start and end in the sorted keys (row = key >> vbits).
DYNG_DEV void mark_row_range(const u64* keys, int64_t i, int64_t N, int vbits, int32_t* row_a, int32_t* row_b) {
  const u64 key = keys[i];
  const u64 r = key >> vbits;
  if (i == 0 || (keys[i - 1] >> vbits) != r) row_a[r] = static_cast<int32_t>(i);
  if (i + 1 == N || (keys[i + 1] >> vbits) != r) row_b[r] = static_cast<int32_t>(i + 1);
}

/// The out-row pass of thread i (every thread of the workgroup calls it: it ends with the
/// workgroup's counter reduction).
DYNG_DEV void out_pairs(const out_pair_params& p, int64_t i) {
  unsigned long long c_ins = 0, c_upd = 0, c_del = 0, c_ign_del = 0, c_ign_ins = 0;
  // weight sum and count-at-maximum deltas of columns 0..3 (named, not an array: a private
  // array indexed by a runtime column would live in scratch memory)
  long long ds0 = 0, ds1 = 0, ds2 = 0, ds3 = 0, dc0 = 0, dc1 = 0, dc2 = 0, dc3 = 0;
  const u64 mask = (1ULL << p.vbits) - 1;
  const int K = p.K;
  if (i == p.N) {
    p.pend_alive[i] = 0;
    p.kill[i] = 0;
  }
  if (i < p.N) {
    const u64 key = p.keys[i];
    if (key != kNoKey) mark_row_range(p.keys, i, p.N, p.vbits, p.row_a, p.row_b);
    if (key == kNoKey) {
      p.pend_alive[i] = 0;
      p.kill[i] = 0;
    } else if (i == 0 || p.keys[i - 1] != key) {
      int64_t j = i + 1;
      while (j < p.N && p.keys[j] == key) ++j;
      const int32_t u = static_cast<int32_t>(key >> p.vbits);
      const int32_t v = static_cast<int32_t>(key & mask);
      const int2 mt = p.meta[u];
      const int32_t* row = p.nbr + mt.x;
      const run_pair rr = find_runs(row, u < p.n_before ? p.base[u] : 0, mt.y, v);
      // the first alive (u,v) of the row in order: the base run's, then the tail's
      auto first_alive = [&]() -> int {
        for (int x = rr.lb; x < rr.ub; ++x) {
          if (row[x] >= 0) return x;
        }
        for (int x = rr.tl; x < rr.tu; ++x) {
          if (row[x] >= 0) return x;
        }
        return -1;
      };
      const int init_hit = u < p.n_before ? first_alive() : -1;
      const bool had = init_hit >= 0;
      // the first operation that overwrote the original first (u,v)'s weights: its recorded old
      // weights are the original ones (else the slot still holds them)
      int64_t first_mod = -1;
      bool structural = false;
      auto account = [&](int k, long long wold, long long wnew, bool removed, bool added) {
        const long long target = p.ctl->col_max[k];
        const long long ds = (added ? wnew : 0) - (removed ? wold : 0);
        const long long dc = ((added && wnew == target) ? 1 : 0) - ((removed && wold == target) ? 1 : 0);
        if (k == 0) {
          ds0 += ds;
          dc0 += dc;
        } else if (k == 1) {
          ds1 += ds;
          dc1 += dc;
        } else if (k == 2) {
          ds2 += ds;
          dc2 += dc;
        } else if (k == 3) {
          ds3 += ds;
          dc3 += dc;
        } else {
          if (ds) atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctl->col_dsum[k]), static_cast<unsigned long long>(ds));
          if (dc) atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctl->col_dcount[k]), static_cast<unsigned long long>(dc));
        }
      };
      for (int64_t q = i; q < j; ++q) {
        const int64_t e = p.vals[q];
        const bool del = p.L.is_del(e);
        p.pend_alive[q] = 0;
        p.kill[q] = 0;
        p.iop_type[e] = kInNone;
        // first alive match: old entries of the runs, then pending entries of this pair
        const int old_hit = first_alive();
        int64_t pend_hit = -1;
        if (old_hit < 0) {
          for (int64_t x = i; x < q; ++x) {
            if (p.pend_alive[x]) { pend_hit = x; break; }
          }
        }
        if (del) {
          if (old_hit >= 0) {
            p.nbr[mt.x + old_hit] = v | kDead;
            if (old_hit < rr.bl) {
              atomicAdd(&p.dead[u], 1);
            } else {
              p.kill[q] = 1;  // a tail entry
            }
            for (int k = 0; k < K; ++k) {
              const int32_t wo = p.w[k * p.slot_cap + mt.x + old_hit];
              p.iop_wold[k * p.N + e] = wo;
              account(k, wo, 0, true, false);
            }
            p.iop_type[e] = kInDelete;
            structural = true;
            ++c_del;
          } else if (pend_hit >= 0) {
            p.pend_alive[pend_hit] = 0;
            for (int k = 0; k < K; ++k) {
              const int32_t wo = p.pend_w[k * p.N + pend_hit];
              p.iop_wold[k * p.N + e] = wo;
              account(k, wo, 0, true, false);
            }
            p.iop_type[e] = kInDelete;
            ++c_del;
          } else {
            ++c_ign_del;
          }
        } else {
          const int64_t bj = p.L.local(e) / p.L.stride;
          if (old_hit >= 0 || pend_hit >= 0) {
            if (!p.upsert) {
              ++c_ign_ins;
            } else {
              if (old_hit >= 0 && old_hit == init_hit && first_mod < 0) first_mod = e;
              for (int k = 0; k < K; ++k) {
                const int32_t wn = p.ins_w[bj * K + k];
                int32_t* slot = old_hit >= 0 ? &p.w[k * p.slot_cap + mt.x + old_hit] : &p.pend_w[k * p.N + pend_hit];
                const int32_t wo = *slot;
                *slot = wn;
                p.iop_wold[k * p.N + e] = wo;
                p.iop_wnew[k * p.N + e] = wn;
                account(k, wo, wn, true, true);
              }
              p.iop_type[e] = kInModify;
              ++c_upd;
            }
          } else {
            p.pend_alive[q] = 1;
            for (int k = 0; k < K; ++k) {
              const int32_t wn = p.ins_w[bj * K + k];
              p.pend_w[k * p.N + q] = wn;
              p.iop_wnew[k * p.N + e] = wn;
              account(k, 0, wn, false, true);
            }
            p.iop_type[e] = kInInsert;
            structural = true;
            ++c_ins;
          }
        }
      }
      // weight-increase classification: final first (u,v) against the original first (u,v)
      if (had) {
        const int32_t* wf = nullptr;
        int64_t wf_stride = 0;
        const int final_hit = first_alive();
        if (final_hit >= 0) {
          wf = p.w + mt.x + final_hit;
          wf_stride = p.slot_cap;
        }
        if (wf == nullptr) {
          for (int64_t x = i; x < j; ++x) {
            if (p.pend_alive[x]) { wf = p.pend_w + x; wf_stride = p.N; break; }
          }
        }
        if (wf != nullptr) {
          for (int64_t q = i; q < j; ++q) {
            const int64_t e = p.vals[q];
            if (p.L.is_del(e)) continue;
            const int64_t loc = p.L.local(e);
            for (int k = 0; k < K; ++k) {
              const int32_t orig = first_mod >= 0 ? p.iop_wold[k * p.N + first_mod] : p.w[k * p.slot_cap + mt.x + init_hit];
              p.e_inc[loc * K + k] = wf[k * wf_stride] > orig ? 1 : 0;
            }
          }
        }
      }
      // simple graphs: the pair's final weights for each of its insertions (what the engine offers
      // the insertion's head, without searching the row again)
      if (p.e_fw != nullptr) {
        const int32_t* wf = nullptr;
        int64_t wf_stride = 0;
        const int final_hit = first_alive();
        if (final_hit >= 0) {
          wf = p.w + mt.x + final_hit;
          wf_stride = p.slot_cap;
        } else {
          for (int64_t x = i; x < j; ++x) {
            if (p.pend_alive[x]) { wf = p.pend_w + x; wf_stride = p.N; break; }
          }
        }
        for (int64_t q = i; q < j; ++q) {
          const int64_t e = p.vals[q];
          if (p.L.is_del(e)) continue;
          const int64_t loc = p.L.local(e);
          for (int k = 0; k < K; ++k) p.e_fw[loc * K + k] = wf != nullptr ? wf[k * wf_stride] : -1;
        }
      }
      const bool first_touch = structural && atomicExch(&p.mark[u], p.mark_gen) != p.mark_gen;
      const int t = wave_append(&p.ctl->touched_out, first_touch);
      if (first_touch) p.touched[t] = u;
    }
  }
  // the workgroup's counters in one pass: wave sums by shuffles, one barrier, then thread 0 adds
  // the waves' sums and bumps the device counters (13 values)
  constexpr int kSums = 13;
  u64 sum[kSums] = {c_ins, c_upd, c_del, c_ign_del, c_ign_ins,
                    static_cast<u64>(ds0), static_cast<u64>(ds1), static_cast<u64>(ds2), static_cast<u64>(ds3),
                    static_cast<u64>(dc0), static_cast<u64>(dc1), static_cast<u64>(dc2), static_cast<u64>(dc3)};
  __shared__ u64 part[kBlock / 64][kSums];
#pragma unroll
  for (int q = 0; q < kSums; ++q) {
    for (int off = 32; off > 0; off >>= 1) sum[q] += static_cast<u64>(__shfl_xor(sum[q], off));
  }
  if (lane() == 0) {
#pragma unroll
    for (int q = 0; q < kSums; ++q) part[threadIdx.x / 64][q] = sum[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < kSums; ++q) {
      u64 t = part[0][q];
      for (int wv = 1; wv < kBlock / 64; ++wv) t += part[wv][q];
      sum[q] = t;
    }
    if (sum[0]) atomicAdd(&p.ctl->inserted_edges, sum[0]);
    if (sum[1]) atomicAdd(&p.ctl->updated_edges, sum[1]);
    if (sum[2]) atomicAdd(&p.ctl->deleted_edges, sum[2]);
    if (sum[3]) atomicAdd(&p.ctl->ignored_deletions, sum[3]);
    if (sum[4]) atomicAdd(&p.ctl->ignored_insertions, sum[4]);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (k >= K) break;
      if (sum[5 + k]) atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctl->col_dsum[k]), sum[5 + k]);
      if (sum[9 + k]) atomicAdd(reinterpret_cast<unsigned long long*>(&p.ctl->col_dcount[k]), sum[9 + k]);
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_out_pairs(out_pair_params p) {
  out_pairs(p, blockIdx.x * int64_t{blockDim.x} + threadIdx.x);
}

struct in_pair_params {
  op_layout L;
  const int32_t* ins_w;  ///< batch insertion weights (j * K + k), for the simple graphs' own replay
  int upsert;            ///< on_existing_insert: 1 upsert, 0 ignore
  const u64* keys;  ///< sorted in-operation keys (v << b | u)
  const uint32_t* vals;  ///< out-operation position of each
  int64_t N;
  int vbits;
  int K;
  const int32_t* iop_type;
  const int32_t* iop_wold;
  const int32_t* iop_wnew;
  int2* meta;
  const int32_t* base;
  int32_t* dead;
  int32_t* nbr;
  int32_t* w;
  int64_t slot_cap;
  int32_t* pend_alive;
  int32_t* kill;
  int32_t* pend_w;
  int32_t* mark;
  int mark_gen;
  int32_t* touched;
  int32_t* row_a;
  int32_t* row_b;
  apply_control* ctl;
};

__global__ void __launch_bounds__(kBlock) k_in_pairs(in_pair_params p) {
  const int64_t i = blockIdx.x * int64_t{blockDim.x} + threadIdx.x;
  if (i == p.N) {
    p.pend_alive[i] = 0;
    p.kill[i] = 0;
  }
  if (i >= p.N) return;
  const u64 key = p.keys[i];
  if (key == kNoKey) {
    p.pend_alive[i] = 0;
    p.kill[i] = 0;
    return;
  }
  mark_row_range(p.keys, i, p.N, p.vbits, p.row_a, p.row_b);
  if (i > 0 && p.keys[i - 1] == key) return;
  int64_t j = i + 1;
  while (j < p.N && p.keys[j] == key) ++j;
  const u64 mask = (1ULL << p.vbits) - 1;
  const int32_t v = static_cast<int32_t>(key >> p.vbits);
  const int32_t u = static_cast<int32_t>(key & mask);
  const int2 mt = p.meta[v];
  const int32_t* row = p.nbr + mt.x;
  const run_pair rr = find_runs(row, p.base[v], mt.y, u);
  const int K = p.K;
  bool structural = false;
  for (int64_t q = i; q < j; ++q) {
    p.pend_alive[q] = 0;
    p.kill[q] = 0;
    const uint32_t o = p.vals[q];  // the operation's batch slot
    const int type = p.iop_type[o];
    if (type == kInNone) continue;  // ignored by the out-row pass
    if (type == kInInsert) {
      p.pend_alive[q] = 1;
      for (int k = 0; k < K; ++k) p.pend_w[k * p.N + q] = p.iop_wnew[k * p.N + o];
      structural = true;
      continue;
    }
    // first alive entry of u with the old weights: the base run's, the tail's, then pending ones
    int old_hit = -1;
    for (int x = rr.lb; x < rr.tu && old_hit < 0; ++x) {
      if (x == rr.ub) x = rr.tl;
      if (x >= rr.tu) break;
      if (row[x] < 0) continue;
      bool same = true;
      for (int k = 0; k < K && same; ++k) same = p.w[k * p.slot_cap + mt.x + x] == p.iop_wold[k * p.N + o];
      if (same) old_hit = x;
    }
    int64_t pend_hit = -1;
    if (old_hit < 0) {
      for (int64_t x = i; x < q && pend_hit < 0; ++x) {
        if (!p.pend_alive[x]) continue;
        bool same = true;
        for (int k = 0; k < K && same; ++k) same = p.pend_w[k * p.N + x] == p.iop_wold[k * p.N + o];
        if (same) pend_hit = x;
      }
    }
    if (old_hit < 0 && pend_hit < 0) {
      atomicOr(&p.ctl->error_bits, static_cast<int>(DYNG_HIP_DEVICE_ERROR_INTERNAL));
      continue;
    }
    if (type == kInDelete) {
      if (old_hit >= 0) {
        p.nbr[mt.x + old_hit] = u | kDead;
        if (old_hit < rr.bl) {
          atomicAdd(&p.dead[v], 1);
        } else {
          p.kill[q] = 1;  // a tail entry
        }
        structural = true;
      } else {
        p.pend_alive[pend_hit] = 0;
      }
    } else {
      for (int k = 0; k < K; ++k) {
        const int32_t wn = p.iop_wnew[k * p.N + o];
        if (old_hit >= 0) p.w[k * p.slot_cap + mt.x + old_hit] = wn; else p.pend_w[k * p.N + pend_hit] = wn;
      }
    }
  }
  const bool first_touch = structural && atomicExch(&p.mark[v], p.mark_gen) != p.mark_gen;
  const int t = wave_append(&p.ctl->touched_in, first_touch);
  if (first_touch) p.touched[t] = v;
}

/// The in-row pass of a simple graph, thread i. A simple graph holds at most one alive (u, v), and
/// its in-row entry (v, u) carries the same weights, so replaying the segment's operations in
/// batch order on the in-row as the out-row pass replays them on the out-row (the first alive old
/// entry, else the first alive pending one; deletions remove it, insertions upsert or append)
/// reaches the same decisions and weights without that pass's results: both passes run in one
/// launch (k_pairs_simple). With parallel edges the copy an operation takes is matched by weight
/// to the out-row's (k_in_pairs, after k_out_pairs).
DYNG_DEV void in_pairs_simple(const in_pair_params& p, int64_t i) {
  if (i == p.N) {
    p.pend_alive[i] = 0;
    p.kill[i] = 0;
  }
  if (i >= p.N) return;
  const u64 key = p.keys[i];
  if (key == kNoKey) {
    p.pend_alive[i] = 0;
    p.kill[i] = 0;
    return;
  }
  mark_row_range(p.keys, i, p.N, p.vbits, p.row_a, p.row_b);
  if (i > 0 && p.keys[i - 1] == key) return;
  int64_t j = i + 1;
  while (j < p.N && p.keys[j] == key) ++j;
  const u64 mask = (1ULL << p.vbits) - 1;
  const int32_t v = static_cast<int32_t>(key >> p.vbits);
  const int32_t u = static_cast<int32_t>(key & mask);
  const int2 mt = p.meta[v];
  const int32_t* row = p.nbr + mt.x;
  const run_pair rr = find_runs(row, p.base[v], mt.y, u);
  const int K = p.K;
  auto first_alive = [&]() -> int {
    for (int x = rr.lb; x < rr.ub; ++x) {
      if (row[x] >= 0) return x;
    }
    for (int x = rr.tl; x < rr.tu; ++x) {
      if (row[x] >= 0) return x;
    }
    return -1;
  };
  bool structural = false;
  for (int64_t q = i; q < j; ++q) {
    p.pend_alive[q] = 0;
    p.kill[q] = 0;
    const int64_t e = p.vals[q];
    const int old_hit = first_alive();
    int64_t pend_hit = -1;
    if (old_hit < 0) {
      for (int64_t x = i; x < q; ++x) {
        if (p.pend_alive[x]) {
          pend_hit = x;
          break;
        }
      }
    }
    if (p.L.is_del(e)) {
      if (old_hit >= 0) {
        p.nbr[mt.x + old_hit] = u | kDead;
        if (old_hit < rr.bl) {
          atomicAdd(&p.dead[v], 1);
        } else {
          p.kill[q] = 1;  // a tail entry
        }
        structural = true;
      } else if (pend_hit >= 0) {
        p.pend_alive[pend_hit] = 0;
      }
    } else {
      const int64_t bj = p.L.local(e) / p.L.stride;
      if (old_hit >= 0 || pend_hit >= 0) {
        if (p.upsert) {
          for (int k = 0; k < K; ++k) {
            const int32_t wn = p.ins_w[bj * K + k];
            if (old_hit >= 0) p.w[k * p.slot_cap + mt.x + old_hit] = wn; else p.pend_w[k * p.N + pend_hit] = wn;
          }
        }
      } else {
        p.pend_alive[q] = 1;
        for (int k = 0; k < K; ++k) p.pend_w[k * p.N + q] = p.ins_w[bj * K + k];
        structural = true;
      }
    }
  }
  const bool first_touch = structural && atomicExch(&p.mark[v], p.mark_gen) != p.mark_gen;
  const int t = wave_append(&p.ctl->touched_in, first_touch);
  if (first_touch) p.touched[t] = v;
}

/// Both pairing passes of a simple graph in one launch: workgroups [0, split) the out-rows,
/// the rest the in-rows (their own replay, independent of the out-row pass).
__global__ void __launch_bounds__(kBlock) k_pairs_simple(out_pair_params po, in_pair_params pi, int split) {
  const int b = static_cast<int>(blockIdx.x);
  if (b < split) {
    out_pairs(po, b * int64_t{blockDim.x} + threadIdx.x);
  } else {
    in_pairs_simple(pi, (b - split) * int64_t{blockDim.x} + threadIdx.x);
  }
}

struct merge_params {
  int2* meta;
  int32_t* cap;
  int32_t* base;
  int32_t* dead;
  int32_t* nbr;
  int32_t* w;
  int64_t slot_cap;
  int K;
  const u64* keys;
  int64_t N;
  int vbits;
  const int32_t* pend_alive;
  const int32_t* pend_prefix;
  const int32_t* kill_prefix;  ///< prefix of the tail entries each operation killed
  const int32_t* pend_w;
  const int32_t* touched;
  const int* touched_count;
  const int32_t* row_a;  ///< the row's operations: keys[row_a[u], row_b[u])
  const int32_t* row_b;
  int32_t* lists;        ///< the merge plan: kModes lists of rows, N + 1 entries each
  int* list_count;       ///< rows per list
  int32_t* stage_nbr;
  int32_t* stage_w;
  int64_t stage_cap;
  long long* stage_used;
  long long* tail;
  long long* garbage;
  unsigned long long* relocated;
  long long* need_reloc;
  long long* need_stage;
  int* error_bits;
  double headroom;
  const int* hold;  ///< null, or a flag that makes the merge do nothing (see k_merge_gate)
};

/// The in-place tail merges of one mode for both stores in one launch: workgroups [0, split)
/// merge out-rows, the others in-rows. The two stores share nothing these merges write (the error
/// bits aside, which they set atomically), so the halves run side by side.
struct merge_sides {
  merge_params side[2];
  int split;
};

/// Runs `body(p, block, blocks)` with this workgroup's store and its index and count among that
/// store's workgroups (one copy of the body: the store's parameters are selected, not branched on).
template <typename Body>
DYNG_DEV void for_side(const merge_sides& m, Body&& body) {
  const int b = static_cast<int>(blockIdx.x);
  const bool in = b >= m.split;
  const merge_params p = in ? m.side[1] : m.side[0];
  if (p.hold != nullptr && *p.hold != 0) return;  // uniform: the gate held the merges back
  body(p, in ? b - m.split : b, in ? static_cast<int>(gridDim.x) - m.split : m.split);
}

constexpr int kMergeBlock = 1024;  ///< threads of a merge workgroup
constexpr int kMergeWaves = kMergeBlock / 64;
constexpr int kMergeSeg = 2048;    ///< operations of one row a workgroup stages in LDS
constexpr int kRegChunks = kMergeSmall / 64;  ///< a short tail held in registers, 64 entries per chunk
constexpr int kRegWeights = 2;                ///< weight columns a tail merge moves per pass
constexpr int kTailMin = 64;                  ///< a tail may always grow to this many entries
constexpr int kTailMax = 4096;                ///< and never past this (a workgroup's registers)
constexpr int kTailPerThread = kTailMax / kMergeBlock;
static_assert(kMergeSmall % 64 == 0, "short tails are whole chunks of 64");
static_assert(kTailMax % kMergeBlock == 0, "a long tail is whole slabs of a workgroup");

constexpr int kLaneTail = 32;  ///< a tail one thread merges in registers: at most this many entries
constexpr int kLaneOps = 8;    ///< and at most this many operations

/// How a touched row is merged: its tail rewritten by one thread (a short tail and a few
/// operations, the common case), by one wave or by a workgroup, or the whole row rewritten as one
/// base run.
enum merge_mode { kTailWave = 0, kTailBlock = 1, kFull = 2, kTailLane = 3, kModes = 4 };

DYNG_DEV int tail_limit(int base_len) { return min(kTailMax, max(kTailMin, base_len >> 3)); }
DYNG_DEV int dead_limit(int base_len) { return max(16, base_len >> 2); }
DYNG_DEV int grown_cap(int len, double headroom) { return len + static_cast<int>(ceil(len * headroom)) + 2; }

/// First index in [0, len) of an LDS array whose value is >= v (strict: > v when `upper`).
DYNG_DEV int lds_bound(const int32_t* a, int len, int32_t v, bool upper) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v || (upper && a[mid] == v)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/// The shape of a touched row's merge.
struct row_plan {
  int mode;
  int base_len, tail_len, new_tail;
  int alive_base, alive_tail, pending;
  int new_len;  ///< full merge: the rewritten row's length
};

DYNG_DEV row_plan plan_row(const merge_params& p, int32_t u, int2 mt, int64_t a, int64_t b) {
  row_plan r;
  r.base_len = p.base[u];
  r.tail_len = mt.y - r.base_len;
  r.pending = p.pend_prefix[b] - p.pend_prefix[a];
  r.alive_tail = r.tail_len - (p.kill_prefix[b] - p.kill_prefix[a]);
  r.alive_base = r.base_len - p.dead[u];
  r.new_tail = r.alive_tail + r.pending;
  r.new_len = r.alive_base + r.new_tail;
  const int nops = static_cast<int>(b - a);
  const bool tail_ok = r.new_tail <= tail_limit(r.base_len) && r.base_len + r.new_tail <= p.cap[u] &&
                       p.dead[u] <= dead_limit(r.base_len);
  if (tail_ok && r.tail_len <= kLaneTail && nops <= kLaneOps) {
    r.mode = kTailLane;
  } else if (tail_ok && r.tail_len <= kMergeSmall && nops <= 64) {
    r.mode = kTailWave;
  } else if (tail_ok && nops <= kMergeSeg) {
    r.mode = kTailBlock;
  } else {
    r.mode = kFull;
  }
  return r;
}

/// Before the merges: each touched row's mode (appended to its list) and, for full merges, the
/// staging slots (its alive entries) and the tail slots (a row that outgrew its capacity), so the
/// host can compact or grow first and no row is refused. One row per thread; the lists are
/// appended per workgroup (one atomic per mode and workgroup: the list counters are single
/// addresses every workgroup hits).
constexpr int kPlanBlock = 1024;

DYNG_DEV void merge_plan(const merge_params& p, int block) {
  __shared__ int wave_offset[kPlanBlock / 64][kModes];
  __shared__ int base[kModes];
  __shared__ long long part[kPlanBlock / 64];
  const int count = *p.touched_count;
  const int t = block * kPlanBlock + threadIdx.x;
  const int w = threadIdx.x / 64;
  long long reloc = 0, stage = 0;
  int mode = -1;
  int32_t u = -1;
  if (t < count) {
    u = p.touched[t];
    const int2 mt = p.meta[u];
    const row_plan r = plan_row(p, u, mt, p.row_a[u], p.row_b[u]);
    mode = r.mode;
    if (r.mode == kFull) {
      stage = r.alive_base + r.alive_tail;
      if (r.new_len > p.cap[u]) reloc = grown_cap(r.new_len, p.headroom);
    }
  }
  int rank = 0;  // among the wave's rows of the same mode
#pragma unroll
  for (int m = 0; m < kModes; ++m) {
    const u64 b = __ballot(mode == m);
    if (mode == m) rank = __popcll(b & lanes_below());
    if (lane() == 0) wave_offset[w][m] = __popcll(b);
  }
  __syncthreads();
  if (threadIdx.x < kModes) {
    const int m = threadIdx.x;
    int run = 0;
    for (int i = 0; i < kPlanBlock / 64; ++i) {
      const int c = wave_offset[i][m];
      wave_offset[i][m] = run;
      run += c;
    }
    base[m] = run > 0 ? atomicAdd(&p.list_count[m], run) : 0;
  }
  __syncthreads();
  if (mode >= 0) p.lists[mode * (p.N + 1) + base[mode] + wave_offset[w][mode] + rank] = u;
  reloc = block_reduce(reloc, part, op_sum{});
  stage = block_reduce(stage, part, op_sum{});
  if (threadIdx.x == 0) {
    if (reloc) atomicAdd(reinterpret_cast<unsigned long long*>(p.need_reloc), static_cast<unsigned long long>(reloc));
    if (stage) atomicAdd(reinterpret_cast<unsigned long long*>(p.need_stage), static_cast<unsigned long long>(stage));
  }
}

__global__ void __launch_bounds__(kPlanBlock) k_merge_plan(merge_sides m) {
  for_side(m, [](const merge_params& p, int block, int) { merge_plan(p, block); });
}

/// Slots a relocating merge leaves free at the arena's tail (the host's plan keeps the same).
constexpr long long kArenaMargin = 1024;

/// After a plan that the host does not read back: whether the merges fit as planned (the
/// arenas' free tails take the relocations, the staging arena the full merges) and no error came
/// first. If not, `hold` makes the merges and the engine run behind them do nothing, and the host
/// plans again after its synchronization. One thread.
__global__ void k_merge_gate(apply_control* c, long long slot_cap_out, long long slot_cap_in, long long stage_cap) {
  const bool fits = c->error_bits == 0 && c->need_reloc_out + kArenaMargin <= slot_cap_out - c->tail_out &&
                    c->need_reloc_in + kArenaMargin <= slot_cap_in - c->tail_in &&
                    c->need_stage_out <= stage_cap && c->need_stage_in <= stage_cap;
  c->hold = fits ? 0 : 1;
}

/// One thread per row whose tail holds at most kLaneTail entries and that has at most kLaneOps
/// operations: the tail and the operations are read into registers with independent loads, every
/// entry's place is counted with static indices (an alive entry: the alive entries before it plus
/// the pending entries with a smaller neighbour; a pending entry: the alive entries up to its
/// neighbour plus the pending entries before it), then each is stored at its place. The weights
/// move KW columns per pass, each pass reading its columns before writing them (KW = 1 for a
/// single weight column: 162 VGPRs instead of 186, three waves per SIMD instead of two). Old
/// entries come first among equal neighbours, as in the wave merge below.
template <int KW>
DYNG_DEV void tail_lane(const merge_params& p, int block, int blocks) {
  const int count = p.list_count[kTailLane];
  const int32_t* list = p.lists + kTailLane * (p.N + 1);
  const u64 mask = (1ULL << p.vbits) - 1;
  const int K = p.K;
  for (int t = block * blockDim.x + threadIdx.x; t < count; t += blocks * blockDim.x) {
    const int32_t u = list[t];
    const int2 mt = p.meta[u];
    const int64_t a = p.row_a[u];
    const int nops = static_cast<int>(p.row_b[u] - a);
    const int bl = p.base[u];
    const int tl = mt.y - bl;
    const int64_t at = mt.x + int64_t{bl};  // the tail's first slot
    int32_t x[kLaneTail];
#pragma unroll
    for (int j = 0; j < kLaneTail; ++j) x[j] = j < tl ? p.nbr[at + j] : -1;
    int32_t ov[kLaneOps];
    bool pend[kLaneOps];
#pragma unroll
    for (int j = 0; j < kLaneOps; ++j) {
      pend[j] = j < nops && p.pend_alive[a + j] != 0;
      ov[j] = pend[j] ? static_cast<int32_t>(p.keys[a + j] & mask) : 0;
    }
    int alive = 0, pending = 0;
    for (int k0 = 0; k0 == 0 || k0 < K; k0 += KW) {
      int32_t wr[KW][kLaneTail];
      int32_t ow[KW][kLaneOps];
#pragma unroll
      for (int j = 0; j < kLaneTail; ++j) {
#pragma unroll
        for (int k = 0; k < KW; ++k) {
          wr[k][j] = (j < tl && k0 + k < K) ? p.w[(k0 + k) * p.slot_cap + at + j] : 0;
        }
      }
#pragma unroll
      for (int j = 0; j < kLaneOps; ++j) {
#pragma unroll
        for (int k = 0; k < KW; ++k) ow[k][j] = (pend[j] && k0 + k < K) ? p.pend_w[(k0 + k) * p.N + a + j] : 0;
      }
      // alive entries, at their places (the tail is sorted, dead entries carry the sign bit)
      alive = 0;
#pragma unroll
      for (int i = 0; i < kLaneTail; ++i) {
        if (x[i] >= 0) {
          int pos = alive;
#pragma unroll
          for (int j = 0; j < kLaneOps; ++j) pos += (pend[j] && ov[j] < x[i]) ? 1 : 0;
          if (k0 == 0) p.nbr[at + pos] = x[i];
#pragma unroll
          for (int k = 0; k < KW; ++k) {
            if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + at + pos] = wr[k][i];
          }
          ++alive;
        }
      }
      // pending entries (the operations are sorted by neighbour, then batch order)
      pending = 0;
#pragma unroll
      for (int j = 0; j < kLaneOps; ++j) {
        if (pend[j]) {
          int pos = pending;
#pragma unroll
          for (int i = 0; i < kLaneTail; ++i) pos += (x[i] >= 0 && x[i] <= ov[j]) ? 1 : 0;
          if (k0 == 0) p.nbr[at + pos] = ov[j];
#pragma unroll
          for (int k = 0; k < KW; ++k) {
            if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + at + pos] = ow[k][j];
          }
          ++pending;
        }
      }
    }
    p.meta[u] = make_int2(mt.x, bl + alive + pending);
  }
}

template <int KW>
__global__ void __launch_bounds__(256) k_tail_lane(merge_sides m) {
  for_side(m, [](const merge_params& p, int block, int blocks) { tail_lane<KW>(p, block, blocks); });
}

/// One wave per row whose tail is short (at most kMergeSmall entries, 64 operations): the tail
/// is read into registers and the merged tail (its alive entries and the pending ones, by
/// neighbour id, old entries first among equal ones) written back in place. An alive entry's
/// place is the alive entries before it plus the pending entries with a smaller neighbour (the
/// operations one per lane); a pending entry's place is the alive entries up to its neighbour
/// (ballots over the registers) plus the pending entries before it.
DYNG_DEV void tail_wave(const merge_params& p, int block, int blocks) {
  const int count = p.list_count[kTailWave];
  const int32_t* list = p.lists + kTailWave * (p.N + 1);
  const u64 mask = (1ULL << p.vbits) - 1;
  const int wave = static_cast<int>((block * int64_t{blockDim.x} + threadIdx.x) / 64);
  const int waves = static_cast<int>((int64_t{blocks} * blockDim.x) / 64);
  const int K = p.K;
  for (int t = wave; t < count; t += waves) {
    const int32_t u = list[t];
    const int2 mt = p.meta[u];
    const int64_t a = p.row_a[u], b = p.row_b[u];
    const int nops = static_cast<int>(b - a);
    const int bl = p.base[u];
    const int tl = mt.y - bl;
    const int64_t at = mt.x + int64_t{bl};  // the tail's first slot
    int32_t x[kRegChunks];
#pragma unroll
    for (int c = 0; c < kRegChunks; ++c) {
      const int jx = c * 64 + lane();
      x[c] = jx < tl ? p.nbr[at + jx] : -1;
    }
    const bool has = lane() < nops;
    const int32_t ov = has ? static_cast<int32_t>(p.keys[a + lane()] & mask) : INT32_MAX;
    const bool opend = has && p.pend_alive[a + lane()] != 0;
    int before[kRegChunks];
#pragma unroll
    for (int c = 0; c < kRegChunks; ++c) before[c] = 0;
    const u64 pend_mask = __ballot(opend);
    for (int j = 0; j < nops; ++j) {
      if (!((pend_mask >> j) & 1)) continue;  // uniform
      const int32_t vj = __shfl(ov, j);
#pragma unroll
      for (int c = 0; c < kRegChunks; ++c) before[c] += (x[c] >= 0 && vj < x[c]) ? 1 : 0;
    }
    // pending entries: their places, computed before any store
    int64_t ppos = -1;
    for (int j = 0; j < nops; ++j) {
      if (!((pend_mask >> j) & 1)) continue;  // uniform
      const int32_t vj = __shfl(ov, j);
      int y = 0;
#pragma unroll
      for (int c = 0; c < kRegChunks; ++c) y += __popcll(__ballot(x[c] >= 0 && x[c] <= vj));
      if (lane() == j) ppos = at + y + __popcll(pend_mask & ((1ULL << j) - 1));
    }
    // the stores, kRegWeights weight columns per pass: every lane of the wave reads its entries'
    // old weights of the pass's columns before any lane writes them
    int carry = 0;
    for (int k0 = 0; k0 == 0 || k0 < K; k0 += kRegWeights) {
      int32_t wr[kRegWeights][kRegChunks];
#pragma unroll
      for (int c = 0; c < kRegChunks; ++c) {
        const int jx = c * 64 + lane();
#pragma unroll
        for (int k = 0; k < kRegWeights; ++k) {
          wr[k][c] = (jx < tl && k0 + k < K) ? p.w[(k0 + k) * p.slot_cap + at + jx] : 0;
        }
      }
      carry = 0;
#pragma unroll
      for (int c = 0; c < kRegChunks; ++c) {
        const bool alive = x[c] >= 0;
        const u64 ballot = __ballot(alive);
        if (alive) {
          const int64_t pos = at + carry + __popcll(ballot & lanes_below()) + before[c];
          if (k0 == 0) p.nbr[pos] = x[c];
#pragma unroll
          for (int k = 0; k < kRegWeights; ++k) {
            if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + pos] = wr[k][c];
          }
        }
        carry += __popcll(ballot);
      }
      if (opend) {
        if (k0 == 0) p.nbr[ppos] = ov;
#pragma unroll
        for (int k = 0; k < kRegWeights; ++k) {
          if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + ppos] = p.pend_w[(k0 + k) * p.N + a + lane()];
        }
      }
    }
    if (lane() == 0) p.meta[u] = make_int2(mt.x, bl + carry + __popcll(pend_mask));
  }
}

__global__ void __launch_bounds__(kMergeBlock) k_tail_wave(merge_sides m) {
  for_side(m, [](const merge_params& p, int block, int blocks) { tail_wave(p, block, blocks); });
}

/// One workgroup per row whose tail needs more than a wave (up to kTailMax entries, kMergeSeg
/// operations): the tail in registers (kTailPerThread per thread), the operations and the
/// alive tail in LDS, the merged tail written back in place. The LDS arrays are the function's
/// own (one set per workgroup, whichever store it merges).
DYNG_DEV void tail_block(const merge_params& p, int block, int blocks) {
  __shared__ int32_t s_v[kMergeSeg];       // neighbour of each operation
  __shared__ int32_t s_pp[kMergeSeg + 1];  // pending entries before each operation
  __shared__ int32_t s_t[kTailMax];        // the alive tail entries, compacted
  __shared__ int s_tot[kTailPerThread * kMergeWaves + 1];
  const int count = p.list_count[kTailBlock];
  const int32_t* list = p.lists + kTailBlock * (p.N + 1);
  const u64 mask = (1ULL << p.vbits) - 1;
  const int tid = static_cast<int>(threadIdx.x);
  const int wid = tid / 64;
  const int K = p.K;
  for (int t = block; t < count; t += blocks) {
    const int32_t u = list[t];
    const int2 mt = p.meta[u];
    const int64_t a = p.row_a[u], b = p.row_b[u];
    const int nops = static_cast<int>(b - a);
    const int bl = p.base[u];
    const int tl = mt.y - bl;
    const int64_t at = mt.x + int64_t{bl};
    const int pa = p.pend_prefix[a];
    for (int j = tid; j < nops; j += kMergeBlock) {
      s_v[j] = static_cast<int32_t>(p.keys[a + j] & mask);
      s_pp[j] = p.pend_prefix[a + j] - pa;
    }
    if (tid == 0) s_pp[nops] = p.pend_prefix[b] - pa;
    int32_t x[kTailPerThread];
    int32_t wr[kRegWeights][kTailPerThread];
    int rank[kTailPerThread];
#pragma unroll
    for (int e = 0; e < kTailPerThread; ++e) {
      const int jx = e * kMergeBlock + tid;
      const bool in = jx < tl;
      x[e] = in ? p.nbr[at + jx] : -1;
#pragma unroll
      for (int k = 0; k < kRegWeights; ++k) wr[k][e] = (in && k < K) ? p.w[k * p.slot_cap + at + jx] : 0;
      const u64 ballot = __ballot(x[e] >= 0);
      rank[e] = __popcll(ballot & lanes_below());
      if (lane() == 0) s_tot[e * kMergeWaves + wid] = __popcll(ballot);
    }
    __syncthreads();
    if (tid == 0) {
      int run = 0;
      for (int i = 0; i < kTailPerThread * kMergeWaves; ++i) {
        const int c = s_tot[i];
        s_tot[i] = run;
        run += c;
      }
      s_tot[kTailPerThread * kMergeWaves] = run;
    }
    __syncthreads();
    const int alive_tail = s_tot[kTailPerThread * kMergeWaves];
    int apos[kTailPerThread];
#pragma unroll
    for (int e = 0; e < kTailPerThread; ++e) {
      apos[e] = s_tot[e * kMergeWaves + wid] + rank[e];
      if (x[e] >= 0) s_t[apos[e]] = x[e];
    }
    __syncthreads();
    // the stores, kRegWeights weight columns per pass (the first pass's weights were read with
    // the tail; a later pass reads its columns, then waits for the whole workgroup before writing)
    for (int k0 = 0; k0 == 0 || k0 < K; k0 += kRegWeights) {
      if (k0 > 0) {
#pragma unroll
        for (int e = 0; e < kTailPerThread; ++e) {
          const int jx = e * kMergeBlock + tid;
#pragma unroll
          for (int k = 0; k < kRegWeights; ++k) {
            wr[k][e] = (jx < tl && k0 + k < K) ? p.w[(k0 + k) * p.slot_cap + at + jx] : 0;
          }
        }
        __syncthreads();
      }
      // pending entries: the alive tail entries up to their neighbour, the pending ones before them
      for (int j = tid; j < nops; j += kMergeBlock) {
        if (s_pp[j + 1] == s_pp[j]) continue;  // not pending
        const int32_t v = s_v[j];
        const int64_t pos = at + lds_bound(s_t, alive_tail, v, true) + s_pp[j];
        if (k0 == 0) p.nbr[pos] = v;
        for (int k = 0; k < kRegWeights; ++k) {
          if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + pos] = p.pend_w[(k0 + k) * p.N + a + j];
        }
      }
#pragma unroll
      for (int e = 0; e < kTailPerThread; ++e) {
        if (x[e] < 0) continue;
        const int64_t pos = at + apos[e] + s_pp[lds_bound(s_v, nops, x[e], false)];
        if (k0 == 0) p.nbr[pos] = x[e];
#pragma unroll
        for (int k = 0; k < kRegWeights; ++k) {
          if (k0 + k < K) p.w[(k0 + k) * p.slot_cap + pos] = wr[k][e];
        }
      }
    }
    if (tid == 0) p.meta[u] = make_int2(mt.x, bl + alive_tail + s_pp[nops]);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kMergeBlock) k_tail_block(merge_sides m) {
  for_side(m, [](const merge_params& p, int block, int blocks) { tail_block(p, block, blocks); });
}

/// Count of the entries of a sorted run in global memory that are < v (or <= v when `upper`).
DYNG_DEV int global_bound(const int32_t* a, int len, int32_t v, bool upper) {
  int lo = 0, hi = len;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int32_t x = a[mid];
    if (x < v || (upper && x == v)) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/// Compact the alive entries of [src, src + len) to dst (a workgroup; returns the count).
DYNG_DEV int compact_alive(const merge_params& p, int64_t src, int len, int64_t dst, int* s_tot) {
  const int tid = static_cast<int>(threadIdx.x);
  const int wid = tid / 64;
  int carry = 0;
  for (int base = 0; base < len; base += kMergeBlock) {
    const int jx = base + tid;
    const int32_t x = jx < len ? p.nbr[src + jx] : -1;
    const u64 ballot = __ballot(x >= 0);
    if (lane() == 0) s_tot[wid] = __popcll(ballot);
    __syncthreads();
    int before = 0, total = 0;
    for (int q = 0; q < kMergeWaves; ++q) {
      before += q < wid ? s_tot[q] : 0;
      total += s_tot[q];
    }
    if (x >= 0) {
      const int64_t pos = dst + carry + before + __popcll(ballot & lanes_below());
      p.stage_nbr[pos] = x;
      for (int k = 0; k < p.K; ++k) p.stage_w[k * p.stage_cap + pos] = p.w[k * p.slot_cap + src + jx];
    }
    carry += total;
    __syncthreads();
  }
  return carry;
}

/// One workgroup per row rewritten whole (its tail or dead entries outgrew their limits, or the
/// row its capacity): the alive base (A) and the alive tail (B) are compacted into the staging
/// arena, then A, B and the pending entries (C) are merged into one base run, in place or at the
/// tail of the arena. An entry's place is its index in its run plus the entries of the other two
/// runs before it: A before B before C among equal neighbours.
__global__ void __launch_bounds__(kMergeBlock) k_full_merge(merge_params p) {
  __shared__ int32_t s_v[kMergeSeg];
  __shared__ int32_t s_pp[kMergeSeg + 1];
  __shared__ int32_t s_b[kTailMax];
  __shared__ int s_tot[kMergeWaves];
  __shared__ long long s_stage, s_dest;
  __shared__ int s_newcap;
  if (p.hold != nullptr && *p.hold != 0) return;  // uniform: the gate held the merges back
  const int count = p.list_count[kFull];
  const int32_t* list = p.lists + kFull * (p.N + 1);
  const u64 mask = (1ULL << p.vbits) - 1;
  const int tid = static_cast<int>(threadIdx.x);
  const int K = p.K;
  for (int t = blockIdx.x; t < count; t += gridDim.x) {
    const int32_t u = list[t];
    const int2 mt = p.meta[u];
    const int64_t a = p.row_a[u], b = p.row_b[u];
    const int nops = static_cast<int>(b - a);
    const row_plan r = plan_row(p, u, mt, a, b);
    if (tid == 0) {
      s_stage = static_cast<long long>(atomicAdd(reinterpret_cast<unsigned long long*>(p.stage_used),
                                                 static_cast<unsigned long long>(r.alive_base + r.alive_tail)));
      s_newcap = p.cap[u];
      s_dest = mt.x;
      if (s_stage + r.alive_base + r.alive_tail > p.stage_cap) {
        atomicOr(p.error_bits, static_cast<int>(DYNG_HIP_DEVICE_ERROR_CAPACITY));
        s_dest = -1;
      } else if (r.new_len > p.cap[u]) {
        const int nc = grown_cap(r.new_len, p.headroom);
        const long long d = static_cast<long long>(
            atomicAdd(reinterpret_cast<unsigned long long*>(p.tail), static_cast<unsigned long long>(nc)));
        if (d + nc > p.slot_cap) {
          atomicOr(p.error_bits, static_cast<int>(DYNG_HIP_DEVICE_ERROR_CAPACITY));
          s_dest = -1;
        } else {
          atomicAdd(reinterpret_cast<unsigned long long*>(p.garbage), static_cast<unsigned long long>(p.cap[u]));
          atomicAdd(p.relocated, 1ULL);
          s_dest = d;
          s_newcap = nc;
        }
      }
    }
    __syncthreads();
    if (s_dest < 0) {
      __syncthreads();
      continue;
    }
    const long long sa = s_stage;
    const long long dest = s_dest;
    const int pa = p.pend_prefix[a];
    const bool ops_lds = nops <= kMergeSeg;
    if (ops_lds) {
      for (int j = tid; j < nops; j += kMergeBlock) {
        s_v[j] = static_cast<int32_t>(p.keys[a + j] & mask);
        s_pp[j] = p.pend_prefix[a + j] - pa;
      }
      if (tid == 0) s_pp[nops] = p.pend_prefix[b] - pa;
    }
    const int A = compact_alive(p, mt.x, r.base_len, sa, s_tot);
    const int B = compact_alive(p, mt.x + int64_t{r.base_len}, r.tail_len, sa + A, s_tot);
    const int32_t* SA = p.stage_nbr + sa;
    const int32_t* SB = p.stage_nbr + sa + A;
    for (int j = tid; j < B; j += kMergeBlock) s_b[j] = SB[j];  // B <= kTailMax
    __syncthreads();  // the staged runs (and the row's old slots are free to overwrite)
    // pending entries before v (strictly smaller neighbour)
    auto pend_before = [&](int32_t v) -> int {
      if (ops_lds) return s_pp[lds_bound(s_v, nops, v, false)];
      const int64_t lt = lower_key(p.keys, a, b, (static_cast<u64>(u) << p.vbits) | static_cast<u64>(v));
      return p.pend_prefix[lt] - pa;
    };
    for (int i = tid; i < A; i += kMergeBlock) {
      const int32_t v = SA[i];
      const int64_t pos = dest + i + lds_bound(s_b, B, v, false) + pend_before(v);
      p.nbr[pos] = v;
      for (int k = 0; k < K; ++k) p.w[k * p.slot_cap + pos] = p.stage_w[k * p.stage_cap + sa + i];
    }
    for (int j = tid; j < B; j += kMergeBlock) {
      const int32_t v = s_b[j];
      const int64_t pos = dest + global_bound(SA, A, v, true) + j + pend_before(v);
      p.nbr[pos] = v;
      for (int k = 0; k < K; ++k) p.w[k * p.slot_cap + pos] = p.stage_w[k * p.stage_cap + sa + A + j];
    }
    for (int64_t q = a + tid; q < b; q += kMergeBlock) {
      if (!p.pend_alive[q]) continue;
      const int32_t v = static_cast<int32_t>(p.keys[q] & mask);
      const int64_t pos = dest + global_bound(SA, A, v, true) + lds_bound(s_b, B, v, true) + (p.pend_prefix[q] - pa);
      p.nbr[pos] = v;
      for (int k = 0; k < K; ++k) p.w[k * p.slot_cap + pos] = p.pend_w[k * p.N + q];
    }
    if (tid == 0) {
      const int len = A + B + r.pending;
      p.meta[u] = make_int2(static_cast<int32_t>(dest), len);
      p.base[u] = len;
      p.dead[u] = 0;
      p.cap[u] = s_newcap;
    }
    __syncthreads();
  }
}

/// Copy one row store into a fresh arena (compaction): one wave per row.
__global__ void k_compact_rows(const int2* meta, const int32_t* nbr, const int32_t* w, int64_t slot_cap,
                               int64_t n, int K, const int64_t* new_off, int2* new_meta, int32_t* new_nbr,
                               int32_t* new_w, int64_t new_slot_cap) {
  const int64_t wave = (blockIdx.x * int64_t{blockDim.x} + threadIdx.x) / 64;
  const int64_t waves = (int64_t{gridDim.x} * blockDim.x) / 64;
  for (int64_t u = wave; u < n; u += waves) {
    const int2 mt = meta[u];
    const int64_t o = new_off[u];
    for (int j = lane(); j < mt.y; j += 64) {
      new_nbr[o + j] = nbr[mt.x + j];
      for (int k = 0; k < K; ++k) new_w[k * new_slot_cap + o + j] = w[k * slot_cap + mt.x + j];
    }
    if (lane() == 0) new_meta[u] = make_int2(static_cast<int32_t>(o), mt.y);
  }
}

/// A freshly built row is one base run.
__global__ void k_base_from_len(const int2* meta, int64_t n, int32_t* base, int32_t* dead) {
  for (int64_t u = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; u < n; u += int64_t{gridDim.x} * blockDim.x) {
    base[u] = meta[u].y;
    dead[u] = 0;
  }
}

__global__ void k_caps_from_len(const int2* meta, int64_t n, double headroom, int32_t* cap, int64_t* cap64) {
  for (int64_t u = blockIdx.x * int64_t{blockDim.x} + threadIdx.x; u < n; u += int64_t{gridDim.x} * blockDim.x) {
    const int len = meta[u].y;
    const int c = len + static_cast<int>(ceil(len * headroom)) + 2;
    cap[u] = c;
    cap64[u] = c;
  }
}

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------

hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

int64_t arena_slots(int64_t rows_slots, int64_t m, int64_t extra) {
  return rows_slots + std::max<int64_t>(m + m / 2, 1 << 16) + extra;
}

void check_slots(int64_t slots) {
  if (slots >= (int64_t{1} << 31) - 1) {
    fail(DYNG_HIP_NOT_SUPPORTED,
         "dyng: the HIP device graph holds at most 2^31 - 1 edge slots per direction (needed " +
             std::to_string(slots) + ")");
  }
}

/// Build one row store from sorted keys (u << b | v) with keep flags and weights.
void build_store(dyng_hip_graph* g, row_store& rs, const u64* keys, const uint32_t* vals, int64_t N,
                 const int32_t* keep, int vbits, int stride, const int32_t* w_in, int from_slots,
                 int64_t in_slot_cap, hipStream_t s) {
  const int64_t n = g->n_cap;
  dbuf<int32_t> keep_prefix;
  keep_prefix.reserve(N + 1);
  exclusive_scan<int32_t>(g->scan_tmp, keep, keep_prefix.p, N + 1, s);
  rs.meta.reserve(n);
  rs.cap.reserve(n);
  rs.base.reserve(n);
  rs.dead.reserve(n);
  fill<int2>(rs.meta.p, n, make_int2(0, 0), s);
  fill<int32_t>(rs.cap.p, n, 0, s);
  fill<int32_t>(rs.base.p, n, 0, s);
  fill<int32_t>(rs.dead.p, n, 0, s);
  dbuf<int64_t> cap64, off64;
  cap64.reserve(g->n + 1);
  off64.reserve(g->n + 1);
  fill<int64_t>(cap64.p, g->n + 1, 0, s);
  hipLaunchKernelGGL(k_row_degrees, dim3(blocks_for(g->n)), dim3(kBlock), 0, s, keys, N, keep_prefix.p,
                     g->n, vbits, g->props.headroom, rs.meta.p, rs.cap.p, cap64.p);
  DYNG_HIP_TRY(hipGetLastError());
  exclusive_scan<int64_t>(g->scan_tmp, cap64.p, off64.p, g->n + 1, s);
  int64_t rows_slots = 0;
  DYNG_HIP_TRY(hipMemcpyAsync(&rows_slots, off64.p + g->n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  hipLaunchKernelGGL(k_set_offsets, dim3(blocks_for(g->n)), dim3(kBlock), 0, s, off64.p, g->n, rs.meta.p);
  const int64_t slots = arena_slots(rows_slots, N, 0);
  check_slots(slots);
  rs.slot_cap = slots;
  rs.tail = rows_slots;
  rs.garbage = 0;
  rs.nbr.reserve(slots);
  rs.w.reserve(slots * g->K);
  hipLaunchKernelGGL(k_fill_rows, dim3(blocks_for(N)), dim3(kBlock), 0, s, keys, vals, N, keep,
                     keep_prefix.p, vbits, stride, rs.meta.p, rs.nbr.p, rs.w.p, rs.slot_cap, g->K, w_in,
                     from_slots, in_slot_cap);
  DYNG_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_base_from_len, dim3(blocks_for(g->n)), dim3(kBlock), 0, s, rs.meta.p, g->n, rs.base.p,
                     rs.dead.p);
  DYNG_HIP_TRY(hipGetLastError());
}

/// In-rows from the out-rows.
void build_in_store(dyng_hip_graph* g, hipStream_t s) {
  const int64_t n = g->n;
  const int vbits = bits_for(std::max<int64_t>(n, 1));
  dbuf<int64_t> len64, pos64;
  len64.reserve(n + 1);
  pos64.reserve(n + 1);
  fill<int64_t>(len64.p, n + 1, 0, s);
  hipLaunchKernelGGL(k_len64, dim3(blocks_for(n)), dim3(kBlock), 0, s, g->out.meta.p, n, len64.p);
  exclusive_scan<int64_t>(g->scan_tmp, len64.p, pos64.p, n + 1, s);
  const int64_t M = g->m;
  dbuf<u64> ka, kb;
  dbuf<uint32_t> va, vb;
  dbuf<int32_t> keep;
  ka.reserve(M);
  kb.reserve(M);
  va.reserve(M);
  vb.reserve(M);
  keep.reserve(M + 1);
  hipLaunchKernelGGL(k_reverse_keys, dim3(blocks_for(n * 64)), dim3(kBlock), 0, s, g->out.meta.p, g->out.nbr.p, n,
                     vbits, ka.p, va.p, pos64.p);
  DYNG_HIP_TRY(hipGetLastError());
  sort_pairs(g->sort_tmp, ka.p, kb.p, va.p, vb.p, M, 2 * vbits, s);
  hipLaunchKernelGGL(k_keep, dim3(blocks_for(M + 1)), dim3(kBlock), 0, s, kb.p, M, 1, keep.p);
  build_store(g, g->in, kb.p, vb.p, M, keep.p, vbits, 1, g->out.w.p, 1, g->out.slot_cap, s);
  DYNG_HIP_TRY(hipStreamSynchronize(s));
}

/// Exact column summaries (max, sum, count of the max) of every objective.
void recompute_columns(dyng_hip_graph* g, hipStream_t s) {
  apply_control* c = g->host_ctl.p;
  std::memset(c, 0, sizeof(apply_control));
  DYNG_HIP_TRY(hipMemcpyAsync(g->ctl.p, c, sizeof(apply_control), hipMemcpyHostToDevice, s));
  // a bounded grid (one reduction per wave, then one atomic per wave and column)
  const unsigned grid = std::min(blocks_for(g->n), 2048u);
  hipLaunchKernelGGL(k_column_summary, dim3(grid), dim3(kBlock), 0, s, g->out.meta.p, g->out.nbr.p, g->out.w.p,
                     g->n, g->out.slot_cap, g->K, 0, g->ctl.p);
  DYNG_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_column_summary, dim3(grid), dim3(kBlock), 0, s, g->out.meta.p, g->out.nbr.p, g->out.w.p,
                     g->n, g->out.slot_cap, g->K, 1, g->ctl.p);
  DYNG_HIP_TRY(hipGetLastError());
  DYNG_HIP_TRY(hipMemcpyAsync(c, g->ctl.p, sizeof(apply_control), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  g->wmax.assign(g->K, 1);
  g->wsum.assign(g->K, 0);
  g->wmax_count.assign(g->K, 0);
  g->wmin.assign(g->K, 1);
  for (int k = 0; k < g->K; ++k) {
    g->wmin[k] = c->col_bad[k] == 0 ? 1 : 0;
    g->wmax[k] = std::max<long long>(1, c->col_max[k]);
    g->wsum[k] = c->col_dsum[k];
    g->wmax_count[k] = c->col_max[k] >= 1 ? c->col_dcount[k] : 0;
  }
}

void compact_store(dyng_hip_graph* g, row_store& rs, int64_t extra, hipStream_t s) {
  const int64_t n = g->n;
  dbuf<int64_t> cap64, off64;
  cap64.reserve(n + 1);
  off64.reserve(n + 1);
  fill<int64_t>(cap64.p, n + 1, 0, s);
  hipLaunchKernelGGL(k_caps_from_len, dim3(blocks_for(n)), dim3(kBlock), 0, s, rs.meta.p, n, g->props.headroom,
                     rs.cap.p, cap64.p);
  exclusive_scan<int64_t>(g->scan_tmp, cap64.p, off64.p, n + 1, s);
  int64_t rows_slots = 0;
  DYNG_HIP_TRY(hipMemcpyAsync(&rows_slots, off64.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  const int64_t slots = arena_slots(rows_slots, g->m, extra);
  check_slots(slots);
  row_store fresh;
  fresh.meta.reserve(g->n_cap);
  fresh.nbr.reserve(slots);
  fresh.w.reserve(slots * g->K);
  fill<int2>(fresh.meta.p, g->n_cap, make_int2(0, 0), s);
  hipLaunchKernelGGL(k_compact_rows, dim3(blocks_for(n * 64)), dim3(kBlock), 0, s, rs.meta.p, rs.nbr.p, rs.w.p,
                     rs.slot_cap, n, g->K, off64.p, fresh.meta.p, fresh.nbr.p, fresh.w.p, slots);
  DYNG_HIP_TRY(hipGetLastError());
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  std::swap(rs.meta.p, fresh.meta.p);
  std::swap(rs.meta.n, fresh.meta.n);
  std::swap(rs.nbr.p, fresh.nbr.p);
  std::swap(rs.nbr.n, fresh.nbr.n);
  std::swap(rs.w.p, fresh.w.p);
  std::swap(rs.w.n, fresh.w.n);
  rs.slot_cap = slots;
  rs.tail = rows_slots;
  rs.garbage = 0;
}

void grow_vertices(dyng_hip_graph* g, int64_t n_after, hipStream_t s) {
  if (n_after <= g->n) return;
  if (n_after > g->n_cap) {
    const int64_t cap = std::max<int64_t>(n_after, g->n_cap + g->n_cap / 2);
    for (row_store* rs : {&g->out, &g->in}) {
      rs->meta.reserve(cap, true, s);
      rs->cap.reserve(cap, true, s);
      rs->base.reserve(cap, true, s);
      rs->dead.reserve(cap, true, s);
    }
    g->mark_out.reserve(cap, true, s);
    g->mark_in.reserve(cap, true, s);
    fill<int32_t>(g->mark_out.p + g->n_cap, cap - g->n_cap, 0, s);
    fill<int32_t>(g->mark_in.p + g->n_cap, cap - g->n_cap, 0, s);
    g->n_cap = cap;
  }
  for (row_store* rs : {&g->out, &g->in}) {
    fill<int2>(rs->meta.p + g->n, n_after - g->n, make_int2(0, 0), s);
    fill<int32_t>(rs->cap.p + g->n, n_after - g->n, 0, s);
    fill<int32_t>(rs->base.p + g->n, n_after - g->n, 0, s);
    fill<int32_t>(rs->dead.p + g->n, n_after - g->n, 0, s);
  }
}

dyng_hip_graph* create_from_edge_list(int device, const dyng_hip_graph_props* props, int64_t n, int64_t count,
                                      int32_t K, const int32_t* src, const int32_t* dst, const int32_t* w,
                                      int32_t from_device, hipStream_t s) {
  if (props == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: graph properties are required");
  if (n < 0 || n >= INT32_MAX) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: the HIP backend takes int32 vertex ids");
  if (K < 1 || K > max_weights) {
    fail(DYNG_HIP_NOT_SUPPORTED, "dyng: the HIP sssp backend takes 1.." + std::to_string(max_weights) +
                                     " weight columns, got " + std::to_string(K));
  }
  if (props->on_existing_insert < 0 || props->on_existing_insert > 2 || props->on_missing_delete < 0 ||
      props->on_missing_delete > 1) {
    fail(DYNG_HIP_INVALID_ARGUMENT,
         "dyng: on_existing_insert must be 0 (upsert), 1 (error) or 2 (ignore) and on_missing_delete 0 "
         "(ignore) or 1 (error)");
  }
  scoped_device guard(device);
  auto g = std::make_unique<dyng_hip_graph>();
  g->device = device;
  g->props = *props;
  if (!(g->props.headroom >= 0.0) || g->props.headroom > 16.0) g->props.headroom = 0.25;
  g->K = K;
  g->n = n;
  g->n_cap = std::max<int64_t>(n, 1);
  const int stride = props->directed ? 1 : 2;
  const int64_t N = count * stride;
  if (N >= (int64_t{1} << 32)) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: too many edges for the HIP device graph");
  const int vbits = bits_for(std::max<int64_t>(n, 1));
  // stage the edge list on the device
  dbuf<int32_t> d_src, d_dst, d_w;
  const int32_t* ps = src;
  const int32_t* pd = dst;
  const int32_t* pw = w;
  if (!from_device) {
    d_src.reserve(count);
    d_dst.reserve(count);
    d_w.reserve(count * K);
    if (count > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(d_src.p, src, count * sizeof(int32_t), hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(d_dst.p, dst, count * sizeof(int32_t), hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(d_w.p, w, count * K * sizeof(int32_t), hipMemcpyHostToDevice, s));
    }
    ps = d_src.p;
    pd = d_dst.p;
    pw = d_w.p;
  }
  g->ctl.reserve(1);
  dbuf<u64> ka, kb;
  dbuf<uint32_t> va, vb;
  dbuf<int32_t> keep;
  ka.reserve(N);
  kb.reserve(N);
  va.reserve(N);
  vb.reserve(N);
  keep.reserve(N + 1);
  if (N > 0) {
    hipLaunchKernelGGL(k_edge_keys, dim3(blocks_for(N)), dim3(kBlock), 0, s, ps, pd, count, stride,
                       props->directed, props->on_self_loop == 1 ? 1 : 0, vbits, ka.p, va.p);
    DYNG_HIP_TRY(hipGetLastError());
    sort_pairs(g->sort_tmp, ka.p, kb.p, va.p, vb.p, N, 2 * vbits, s);
  }
  hipLaunchKernelGGL(k_keep, dim3(blocks_for(N + 1)), dim3(kBlock), 0, s, kb.p, N, props->allow_parallel, keep.p);
  DYNG_HIP_TRY(hipGetLastError());
  build_store(g.get(), g->out, kb.p, vb.p, N, keep.p, vbits, stride, pw, 0, 0, s);
  // edges kept
  {
    dbuf<int32_t> kp;
    kp.reserve(N + 1);
    exclusive_scan<int32_t>(g->scan_tmp, keep.p, kp.p, N + 1, s);
    int32_t mm = 0;
    DYNG_HIP_TRY(hipMemcpyAsync(&mm, kp.p + N, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    g->m = mm;
  }
  g->mark_out.reserve(g->n_cap);
  g->mark_in.reserve(g->n_cap);
  fill<int32_t>(g->mark_out.p, g->n_cap, 0, s);
  fill<int32_t>(g->mark_in.p, g->n_cap, 0, s);
  build_in_store(g.get(), s);
  recompute_columns(g.get(), s);
  static uint64_t next_state = 1;
  g->state_id = __atomic_fetch_add(&next_state, 1, __ATOMIC_RELAXED);
  return g.release();
}

void copy_rows_to_host(const row_store& rs, int64_t n, int K, int64_t* row_ptr, int32_t* col, int32_t* w) {
  std::vector<int2> meta(static_cast<size_t>(n));
  std::vector<int32_t> base(static_cast<size_t>(n));
  if (n > 0) {
    DYNG_HIP_TRY(hipMemcpy(meta.data(), rs.meta.p, n * sizeof(int2), hipMemcpyDeviceToHost));
    DYNG_HIP_TRY(hipMemcpy(base.data(), rs.base.p, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  std::vector<int32_t> nbr(static_cast<size_t>(rs.slot_cap));
  std::vector<int32_t> wt(static_cast<size_t>(rs.slot_cap) * K);
  if (rs.slot_cap > 0) {
    DYNG_HIP_TRY(hipMemcpy(nbr.data(), rs.nbr.p, rs.slot_cap * sizeof(int32_t), hipMemcpyDeviceToHost));
    DYNG_HIP_TRY(hipMemcpy(wt.data(), rs.w.p, rs.slot_cap * K * sizeof(int32_t), hipMemcpyDeviceToHost));
  }
  // each row in the reference's order: its base and tail runs merged by neighbour id (base
  // entries first among equal ones), dead entries dropped
  row_ptr[0] = 0;
  for (int64_t u = 0; u < n; ++u) {
    int64_t alive = 0;
    for (int j = 0; j < meta[u].y; ++j) alive += nbr[meta[u].x + j] >= 0 ? 1 : 0;
    row_ptr[u + 1] = row_ptr[u] + alive;
  }
  const int64_t m = row_ptr[n];
  for (int64_t u = 0; u < n; ++u) {
    const int64_t off = meta[u].x;
    int i = 0, j = base[u];
    const int bend = base[u], tend = meta[u].y;
    int64_t o = row_ptr[u];
    auto put = [&](int x) {
      col[o] = nbr[off + x];
      for (int k = 0; k < K; ++k) w[k * m + o] = wt[k * rs.slot_cap + off + x];
      ++o;
    };
    for (;;) {
      while (i < bend && nbr[off + i] < 0) ++i;
      while (j < tend && nbr[off + j] < 0) ++j;
      if (i == bend && j == tend) break;
      if (j == tend || (i < bend && nbr[off + i] <= nbr[off + j])) {
        put(i++);
      } else {
        put(j++);
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// The apply (host orchestration)
// ---------------------------------------------------------------------------------------------

namespace {

/// Whether the apply times its stages: DYNG_HIP_STAGE_TIMES=1 in the environment. Each timed
/// event costs the stream a few microseconds, so the stages are not timed by default.
bool stage_timing() {
  static const bool on = [] {
    const char* v = std::getenv("DYNG_HIP_STAGE_TIMES");
    return v != nullptr && v[0] == '1';
  }();
  return on;
}

void mark_stage(dyng_hip_graph* g, int i, hipStream_t s) {
  if (!stage_timing()) return;
  if (g->apply_ev[i] == nullptr) DYNG_HIP_TRY(hipEventCreate(&g->apply_ev[i]));
  DYNG_HIP_TRY(hipEventRecord(g->apply_ev[i], s));
}

/// Stage durations from the recorded events (after the stream has been synchronized).
void stage_times(dyng_hip_graph* g, bool rows, dyng_hip_apply_summary* out) {
  for (double& t : out->stage_ms) t = 0.0;
  if (!stage_timing()) return;
  float ms = 0.0f;
  if (hipEventElapsedTime(&ms, g->apply_ev[0], g->apply_ev[1]) == hipSuccess) out->stage_ms[0] = ms;
  if (!rows) return;
  for (int i = 2; i < 12; ++i) {
    if (hipEventElapsedTime(&ms, g->apply_ev[i], g->apply_ev[i + 1]) == hipSuccess) out->stage_ms[i - 1] = ms;
  }
}

}  // namespace

void graph_grow(dyng_hip_graph* g, int64_t n, hipStream_t s) {
  if (n <= g->n) return;
  if (n >= INT32_MAX) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: the HIP backend takes int32 vertex ids");
  grow_vertices(g, n, s);
  g->n = n;
}

void scan_batch(dyng_hip_graph* g, const dyng_hip_batch* b, hipStream_t s, batch_scan& sc) {
  if (b == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: graph apply: no batch");
  mark_stage(g, 0, s);
  const int K = g->K;
  const int64_t num_ins = b->num_ins, num_del = b->num_del;
  if (num_ins < 0 || num_del < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: negative batch sizes");
  if (num_ins > 0 && b->num_weights != K) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: the batch has " + std::to_string(b->num_weights) +
                                        " weights per insertion but the graph has " + std::to_string(K));
  }
  // stage host batches on the device
  const int32_t *ins_src = b->ins_src, *ins_dst = b->ins_dst, *ins_w = b->ins_w, *del_src = b->del_src,
                *del_dst = b->del_dst;
  if (!b->on_device) {
    g->b_ins_src.reserve(num_ins);
    g->b_ins_dst.reserve(num_ins);
    g->b_ins_w.reserve(num_ins * K);
    g->b_del_src.reserve(num_del);
    g->b_del_dst.reserve(num_del);
    if (num_ins > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(g->b_ins_src.p, ins_src, num_ins * 4, hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(g->b_ins_dst.p, ins_dst, num_ins * 4, hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(g->b_ins_w.p, ins_w, num_ins * K * 4, hipMemcpyHostToDevice, s));
    }
    if (num_del > 0) {
      DYNG_HIP_TRY(hipMemcpyAsync(g->b_del_src.p, del_src, num_del * 4, hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(g->b_del_dst.p, del_dst, num_del * 4, hipMemcpyHostToDevice, s));
    }
    ins_src = g->b_ins_src.p;
    ins_dst = g->b_ins_dst.p;
    ins_w = g->b_ins_w.p;
    del_src = g->b_del_src.p;
    del_dst = g->b_del_dst.p;
  }
  apply_control* hc = g->host_ctl.p;
  // ---- scan the batch (ids, self-loops, weights) before anything changes ----
  std::memset(hc, 0, sizeof(apply_control));
  hc->min_id = INT32_MAX;
  hc->max_id = -1;
  hc->min_weight = INT32_MAX;
  for (int k = 0; k < max_weights; ++k) hc->batch_wmin[k] = 1;
  DYNG_HIP_TRY(hipMemcpyAsync(g->ctl.p, hc, sizeof(apply_control), hipMemcpyHostToDevice, s));
  if (num_ins + num_del > 0) {
    hipLaunchKernelGGL(k_batch_scan, dim3(std::min<unsigned>(blocks_for(std::max(num_ins, num_del), kScanBlock), 64)),
                       dim3(kScanBlock), 0, s, ins_src, ins_dst, ins_w, num_ins, del_src, del_dst, num_del, K, g->ctl.p);
    DYNG_HIP_TRY(hipGetLastError());
  }
  mark_stage(g, 1, s);
  DYNG_HIP_TRY(hipMemcpyAsync(hc, g->ctl.p, sizeof(apply_control), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (hc->min_id < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: the batch names a negative vertex id");
  if (g->props.on_self_loop == 2 && hc->self_loops_ins + hc->self_loops_del > 0) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: the batch has a self-loop and on_self_loop is error");
  }
  int64_t n_after = g->n;
  if (hc->max_id >= 0) {
    if (!g->props.allow_vertex_growth && hc->max_id >= g->n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: an insertion names a vertex >= " + std::to_string(g->n) +
                                          " and allow_vertex_growth is false");
    }
    if (hc->max_id >= INT32_MAX - 1) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: an insertion names the largest vertex id");
    // a self-loop insertion that is dropped does not grow the vertex set; keep the rule simple by
    // growing only for kept insertions (the scan covers them all; drop is rare)
    n_after = std::max<int64_t>(n_after, static_cast<int64_t>(hc->max_id) + 1);
  }
  sc.n_after = n_after;
  sc.num_ins = num_ins;
  sc.num_del = num_del;
  sc.ins_src = ins_src;
  sc.ins_dst = ins_dst;
  sc.ins_w = ins_w;
  sc.del_src = del_src;
  sc.del_dst = del_dst;
  sc.wmax.assign(K, 0);
  sc.wmin.assign(K, 1);
  for (int k = 0; k < K; ++k) {
    sc.wmax[k] = num_ins > 0 ? hc->batch_wmax[k] : 0;
    sc.wmin[k] = hc->batch_wmin[k];
  }
}

/// The merge parameters of one row store for a batch of N operations (keys of `vbits` bits).
merge_params merge_setup(dyng_hip_graph* g, bool outs, int64_t N, int vbits, const int* hold) {
  apply_control* dc = g->ctl.p;
  row_store& rs = outs ? g->out : g->in;
  merge_params mp{};
  mp.meta = rs.meta.p;
  mp.cap = rs.cap.p;
  mp.base = rs.base.p;
  mp.dead = rs.dead.p;
  mp.nbr = rs.nbr.p;
  mp.w = rs.w.p;
  mp.slot_cap = rs.slot_cap;
  mp.K = g->K;
  mp.keys = outs ? g->keys_b.p : g->keys_b.p + N;
  mp.N = N;
  mp.vbits = vbits;
  mp.pend_alive = outs ? g->pend_alive.p : g->ipend_alive.p;
  mp.pend_prefix = outs ? g->pend_prefix.p : g->ipend_prefix.p;
  mp.kill_prefix = outs ? g->kill_prefix.p : g->ikill_prefix.p;
  mp.pend_w = outs ? g->pend_w.p : g->ipend_w.p;
  mp.touched = outs ? g->touched_out.p : g->touched_in.p;
  mp.touched_count = outs ? &dc->touched_out : &dc->touched_in;
  mp.row_a = outs ? g->row_a_out.p : g->row_a_in.p;
  mp.row_b = outs ? g->row_b_out.p : g->row_b_in.p;
  mp.lists = g->merge_lists.p + (outs ? 0 : kModes * (N + 1));
  mp.list_count = outs ? dc->modes_out : dc->modes_in;
  mp.stage_nbr = g->stage_nbr.p;
  mp.stage_w = g->stage_w.p;
  mp.stage_cap = g->stage_cap;
  mp.stage_used = outs ? &dc->stage_out : &dc->stage_in;
  mp.tail = outs ? &dc->tail_out : &dc->tail_in;
  mp.garbage = outs ? &dc->garbage_out : &dc->garbage_in;
  mp.relocated = &dc->relocated_rows;
  mp.need_reloc = outs ? &dc->need_reloc_out : &dc->need_reloc_in;
  mp.need_stage = outs ? &dc->need_stage_out : &dc->need_stage_in;
  mp.error_bits = &dc->error_bits;
  mp.headroom = g->props.headroom;
  mp.hold = hold;
  return mp;
}

unsigned plan_groups(int64_t N) { return static_cast<unsigned>((N + kPlanBlock - 1) / kPlanBlock); }  // rows <= N

/// Launch the tail merges of one mode for both stores: `groups` workgroups per store.
void launch_tails(void (*kernel)(merge_sides), merge_sides ms, int out_groups, int in_groups, int block,
                  hipStream_t s) {
  if (out_groups + in_groups == 0) return;
  ms.split = out_groups;
  hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(out_groups + in_groups)), dim3(block), 0, s, ms);
}

/// The merge plan and the merges with the plan read back first: the launches sized from it, the
/// arenas compacted and the staging arena grown as it needs. `replan`: a plan already ran (its
/// counts are cleared first).
void merges_exact(dyng_hip_graph* g, int64_t N, int vbits, hipStream_t s, dyng_hip_apply_summary* out, bool replan) {
  apply_control* hc = g->host_ctl.p;
  apply_control* dc = g->ctl.p;
  const int K = g->K;
  for (int attempt = 0;; ++attempt) {
    if (attempt > 0 || replan) {
      // planning again: clear what the last plan counted (the first plan of a batch finds them
      // zero, as the control block was uploaded zeroed at the start of the apply)
      hc->need_reloc_out = hc->need_reloc_in = hc->need_stage_out = hc->need_stage_in = 0;
      for (int i = 0; i < kModes; ++i) hc->modes_out[i] = hc->modes_in[i] = 0;
      DYNG_HIP_TRY(hipMemcpyAsync(&dc->need_reloc_out, &hc->need_reloc_out, 4 * sizeof(long long),
                                  hipMemcpyHostToDevice, s));
      DYNG_HIP_TRY(hipMemcpyAsync(dc->modes_out, hc->modes_out, 2 * sizeof(hc->modes_out), hipMemcpyHostToDevice,
                                  s));
    }
    merge_sides plan{};  // both stores' plans in one launch: out-rows first
    plan.side[0] = merge_setup(g, true, N, vbits, nullptr);
    plan.side[1] = merge_setup(g, false, N, vbits, nullptr);
    plan.split = static_cast<int>(plan_groups(N));
    hipLaunchKernelGGL(k_merge_plan, dim3(2 * plan_groups(N)), dim3(kPlanBlock), 0, s, plan);
    DYNG_HIP_TRY(hipGetLastError());
    DYNG_HIP_TRY(hipMemcpyAsync(hc, dc, sizeof(apply_control), hipMemcpyDeviceToHost, s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    if (out) out->host_plans += 1;
    if (hc->error_bits != 0) return;  // reported by the caller
    bool compacted = false;
    for (int side = 0; side < 2; ++side) {
      row_store& rs = side == 0 ? g->out : g->in;
      const int64_t need = (side == 0 ? hc->need_reloc_out : hc->need_reloc_in) + kArenaMargin;
      if (rs.slot_cap - rs.tail < need) {
        if (attempt > 2) fail(DYNG_HIP_RUNTIME_ERROR, "dyng: graph apply: the row arena cannot take the batch");
        compact_store(g, rs, 2 * need, s);
        if (out) out->compactions += 1;
        compacted = true;
      }
    }
    if (!compacted) break;
    // compaction recomputed the capacities: plan again on the fresh arenas
    hc->tail_out = g->out.tail;
    hc->tail_in = g->in.tail;
    hc->garbage_out = hc->garbage_in = 0;
    DYNG_HIP_TRY(hipMemcpyAsync(&dc->tail_out, &hc->tail_out, 2 * sizeof(long long), hipMemcpyHostToDevice, s));
    DYNG_HIP_TRY(hipMemcpyAsync(&dc->garbage_out, &hc->garbage_out, 2 * sizeof(long long), hipMemcpyHostToDevice, s));
  }
  const int64_t stage_need = std::max(hc->need_stage_out, hc->need_stage_in);
  if (stage_need > g->stage_cap) {
    g->stage_cap = stage_need + stage_need / 4 + 4096;
    g->stage_nbr.reserve(g->stage_cap);
    g->stage_w.reserve(g->stage_cap * K);
  }
  mark_stage(g, 10, s);
  // the in-place tail merges: one launch per mode for both stores (out-rows in the first
  // workgroups, in-rows in the rest), grids from the plan
  merge_sides ms{};
  ms.side[0] = merge_setup(g, true, N, vbits, nullptr);
  ms.side[1] = merge_setup(g, false, N, vbits, nullptr);
  auto groups = [](int rows, int per_group, int cap) { return std::min((rows + per_group - 1) / per_group, cap); };
  launch_tails(g->K == 1 ? k_tail_lane<1> : k_tail_lane<kRegWeights>, ms, groups(hc->modes_out[kTailLane], 256, 8192),
               groups(hc->modes_in[kTailLane], 256, 8192), 256, s);
  launch_tails(k_tail_wave, ms, groups(hc->modes_out[kTailWave], kMergeWaves, 8192),
               groups(hc->modes_in[kTailWave], kMergeWaves, 8192), kMergeBlock, s);
  launch_tails(k_tail_block, ms, groups(hc->modes_out[kTailBlock], 1, 4096), groups(hc->modes_in[kTailBlock], 1, 4096),
               kMergeBlock, s);
  DYNG_HIP_TRY(hipGetLastError());
  mark_stage(g, 11, s);
  // whole rows through the staging arena, which both stores use: one store at a time
  for (int side = 0; side < 2; ++side) {
    const int rows = side == 0 ? hc->modes_out[kFull] : hc->modes_in[kFull];
    if (rows > 0) {
      hipLaunchKernelGGL(k_full_merge, dim3(static_cast<unsigned>(std::min(rows, 4096))), dim3(kMergeBlock), 0, s,
                         ms.side[side]);
    }
  }
  DYNG_HIP_TRY(hipGetLastError());
  mark_stage(g, 12, s);
}

/// The merge plan and the merges queued together for a small batch, with no read-back in between:
/// a gate on the device checks that the arenas take what the plan asks (k_merge_gate) and holds
/// the merges back otherwise; the grids are sized for the batch (every touched row is one of its
/// at most N + 1 rows per store) and the merges loop over the rows the plan listed.
void merges_speculative(dyng_hip_graph* g, int64_t N, int vbits, hipStream_t s) {
  apply_control* dc = g->ctl.p;
  const int* hold = &dc->hold;
  merge_sides plan{};
  plan.side[0] = merge_setup(g, true, N, vbits, nullptr);
  plan.side[1] = merge_setup(g, false, N, vbits, nullptr);
  plan.split = static_cast<int>(plan_groups(N));
  hipLaunchKernelGGL(k_merge_plan, dim3(2 * plan_groups(N)), dim3(kPlanBlock), 0, s, plan);
  hipLaunchKernelGGL(k_merge_gate, dim3(1), dim3(1), 0, s, dc, static_cast<long long>(g->out.slot_cap),
                     static_cast<long long>(g->in.slot_cap), static_cast<long long>(g->stage_cap));
  DYNG_HIP_TRY(hipGetLastError());
  mark_stage(g, 10, s);
  merge_sides ms{};
  ms.side[0] = merge_setup(g, true, N, vbits, hold);
  ms.side[1] = merge_setup(g, false, N, vbits, hold);
  const int rows = static_cast<int>(N + 1);
  const int lane = std::min((rows + 255) / 256, 1024);
  const int wave = std::min((rows + kMergeWaves - 1) / kMergeWaves, 512);
  const int block = std::min(rows, 256);
  launch_tails(g->K == 1 ? k_tail_lane<1> : k_tail_lane<kRegWeights>, ms, lane, lane, 256, s);
  launch_tails(k_tail_wave, ms, wave, wave, kMergeBlock, s);
  launch_tails(k_tail_block, ms, block, block, kMergeBlock, s);
  DYNG_HIP_TRY(hipGetLastError());
  mark_stage(g, 11, s);
  for (int side = 0; side < 2; ++side) {
    hipLaunchKernelGGL(k_full_merge, dim3(static_cast<unsigned>(block)), dim3(kMergeBlock), 0, s, ms.side[side]);
  }
  DYNG_HIP_TRY(hipGetLastError());
  mark_stage(g, 12, s);
}

/// Throw the invalid-argument error of the offending operation of rank `rank` (k_check_pairs):
/// the message names the operation's kind, batch index and (u, v), as apply_host.cpp:117-120 and
/// 222-225 do. Nothing has changed yet.
[[noreturn]] void report_batch_error(dyng_hip_graph* g, const batch_scan& sc, const op_layout& L,
                                     unsigned long long rank, hipStream_t s) {
  const bool row_phase = (rank >> 63) != 0;
  const int64_t j = static_cast<int64_t>(rank & 0x7fffffffULL);
  const bool second = ((rank >> 31) & 1ULL) != 0;
  const bool del = !row_phase || (second ? L.first_is_del == 0 : L.first_is_del != 0);
  int32_t uv[2] = {0, 0};
  const int32_t* src = del ? sc.del_src : sc.ins_src;
  const int32_t* dst = del ? sc.del_dst : sc.ins_dst;
  DYNG_HIP_TRY(hipMemcpyAsync(&uv[0], src + j, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipMemcpyAsync(&uv[1], dst + j, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (row_phase) {
    // the arc of the row: (v, u) for the reverse arc of an undirected edge
    const int32_t row = static_cast<int32_t>((rank >> 32) & 0x7fffffffULL);
    if (row != uv[0]) std::swap(uv[0], uv[1]);
  }
  (void)g;
  const std::string op = std::string(del ? "deletion " : "insertion ") + std::to_string(j) + " (" +
                         std::to_string(uv[0]) + ", " + std::to_string(uv[1]) + ")";
  throw batch_refused(DYNG_HIP_INVALID_ARGUMENT,
                      "dyng: " + op +
                          (del ? " names a missing edge and batch_semantics::on_missing_delete is error"
                               : " names an existing edge and batch_semantics::on_existing_insert is error"));
}

apply_pending apply_launch(dyng_hip_graph* g, const batch_scan& sc, hipStream_t s, dyng_hip_apply_summary* out,
                           hipEvent_t wait) {
  const int K = g->K;
  // G_t as the batch found it (the repair of a result one apply behind reads it)
  g->last.valid = false;
  g->last.max_weight.assign(K, 1);
  g->last.batch_wmin.assign(K, 1);
  g->last.wsum.assign(K, 0);
  for (int k = 0; k < K; ++k) {
    g->last.max_weight[k] = std::max<int64_t>(std::max<int64_t>(sc.wmax[k], 1), g->wmax[k]);
    g->last.batch_wmin[k] = sc.wmin[k];
    g->last.wsum[k] = g->wsum[k];
  }
  g->last.m = g->m;
  g->last.n = g->n;
  g->last.num_ins = sc.num_ins;
  const int64_t num_ins = sc.num_ins, num_del = sc.num_del;
  const int32_t *ins_src = sc.ins_src, *ins_dst = sc.ins_dst, *ins_w = sc.ins_w, *del_src = sc.del_src,
                *del_dst = sc.del_dst;
  const int64_t n_after = sc.n_after;
  apply_control* hc = g->host_ctl.p;
  const int64_t n_before = g->n;
  const int stride = g->props.directed ? 1 : 2;
  const int64_t N = (num_ins + num_del) * stride;
  if (N >= (int64_t{1} << 31)) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: batch too large");
  grow_vertices(g, n_after, s);
  // deletions below the old vertex count that name vertices of the growth are fine; ids in the
  // batch >= n_after only occur in deletions (handled in Step 0)
  const int vbits = bits_for(std::max<int64_t>(n_after, 1));
  apply_pending ap;
  ap.n_before = n_before;
  ap.N = N;
  ap.vbits = vbits;
  ap.wait = wait;
  op_layout L;
  L.first_is_del = g->props.deletions_first ? 1 : 0;
  L.num_first = L.first_is_del ? num_del : num_ins;
  L.num_second = L.first_is_del ? num_ins : num_del;
  L.stride = stride;
  g->num_del_slots = num_del * stride;
  g->num_ins_slots = num_ins * stride;
  g->del_src.reserve(g->num_del_slots);
  g->del_dst.reserve(g->num_del_slots);
  g->ins_src.reserve(g->num_ins_slots);
  g->ins_dst.reserve(g->num_ins_slots);
  g->ins_inc.reserve(g->num_ins_slots * K);
  if (!g->props.allow_parallel) g->ins_fw.reserve(g->num_ins_slots * K);
  g->keys_a.reserve(2 * N);  // out-row keys, then in-row keys
  g->keys_b.reserve(2 * N);
  g->vals_a.reserve(2 * N);
  g->vals_b.reserve(2 * N);
  g->pend_alive.reserve(N + 1);
  g->kill.reserve(N + 1);
  g->pend_prefix.reserve(N + 1);
  g->kill_prefix.reserve(N + 1);
  g->pend_w.reserve(N * K);
  g->iop_type.reserve(N);
  g->iop_wold.reserve(N * K);
  g->iop_wnew.reserve(N * K);
  g->ipend_alive.reserve(N + 1);
  g->ikill.reserve(N + 1);
  g->ipend_prefix.reserve(N + 1);
  g->ikill_prefix.reserve(N + 1);
  g->ipend_w.reserve(N * K);
  g->touched_out.reserve(N + 1);
  g->touched_in.reserve(N + 1);
  g->row_a_out.reserve(g->n_cap);
  g->row_b_out.reserve(g->n_cap);
  g->row_a_in.reserve(g->n_cap);
  g->row_b_in.reserve(g->n_cap);
  if (++g->mark_generation == INT32_MAX) {
    fill<int32_t>(g->mark_out.p, g->n_cap, 0, s);
    fill<int32_t>(g->mark_in.p, g->n_cap, 0, s);
    g->mark_generation = 1;
  }
  // ---- control block: counters, tails, the column targets ----
  std::memset(hc, 0, sizeof(apply_control));
  hc->first_error = kNoError;
  for (int k = 0; k < K; ++k) hc->col_max[k] = std::max<int64_t>(g->wmax[k], sc.wmax[k]);
  hc->tail_out = g->out.tail;
  hc->tail_in = g->in.tail;
  DYNG_HIP_TRY(hipMemcpyAsync(g->ctl.p, hc, sizeof(apply_control), hipMemcpyHostToDevice, s));
  apply_control* dc = g->ctl.p;
  if (N > 0) {
    mark_stage(g, 2, s);
    hipLaunchKernelGGL(k_make_ops, dim3(blocks_for(N)), dim3(kBlock), 0, s, L, ins_src, ins_dst, del_src, del_dst,
                       g->props.directed, g->props.on_self_loop == 1 ? 1 : 0, n_after, vbits, g->keys_a.p,
                       g->vals_a.p, g->del_src.p, g->del_dst.p, g->ins_src.p, g->ins_dst.p, g->ins_inc.p,
                       g->props.allow_parallel ? nullptr : g->ins_fw.p, K, g->props.on_missing_delete == 1 ? 1 : 0,
                       dc);
    DYNG_HIP_TRY(hipGetLastError());
    mark_stage(g, 3, s);
    // one stable sort of both directions' keys (2N of them, 2 * vbits + 1 bits): one sort of 2N
    // keys costs little more than one of N, since rocPRIM's small-size path is launch-bound
    sort_pairs(g->sort_tmp, g->keys_a.p, g->keys_b.p, g->vals_a.p, g->vals_b.p, 2 * N, 2 * vbits + 1, s);
    hipLaunchKernelGGL(k_untag, dim3(std::min<unsigned>(blocks_for(2 * N), 1024)), dim3(kBlock), 0, s, g->keys_b.p,
                       2 * N, vbits);
    if (g->props.on_missing_delete == 1 || g->props.on_existing_insert == 1) {
      // the error modes: every operation checked before anything changes (one more launch and
      // synchronization, only in these modes)
      hipLaunchKernelGGL(k_check_pairs, dim3(blocks_for(N)), dim3(kBlock), 0, s, L, g->keys_b.p, g->vals_b.p, N,
                         vbits, n_before, g->out.meta.p, g->out.base.p, g->out.nbr.p,
                         g->props.on_missing_delete == 1 ? 1 : 0, g->props.on_existing_insert == 1 ? 1 : 0, dc);
      DYNG_HIP_TRY(hipGetLastError());
      DYNG_HIP_TRY(hipMemcpyAsync(&hc->first_error, &dc->first_error, sizeof(hc->first_error),
                                  hipMemcpyDeviceToHost, s));
      DYNG_HIP_TRY(hipStreamSynchronize(s));
      if (hc->first_error != kNoError) report_batch_error(g, sc, L, hc->first_error, s);
    }
    mark_stage(g, 4, s);
    out_pair_params op{};
    op.L = L;
    op.keys = g->keys_b.p;
    op.vals = g->vals_b.p;
    op.N = N;
    op.vbits = vbits;
    op.K = K;
    op.n_before = n_before;
    op.upsert = g->props.on_existing_insert == 0 ? 1 : 0;
    op.ins_w = ins_w;
    op.meta = g->out.meta.p;
    op.base = g->out.base.p;
    op.dead = g->out.dead.p;
    op.nbr = g->out.nbr.p;
    op.w = g->out.w.p;
    op.slot_cap = g->out.slot_cap;
    op.pend_alive = g->pend_alive.p;
    op.kill = g->kill.p;
    op.pend_w = g->pend_w.p;
    op.e_inc = g->ins_inc.p;
    op.e_fw = g->props.allow_parallel ? nullptr : g->ins_fw.p;
    op.iop_type = g->iop_type.p;
    op.iop_wold = g->iop_wold.p;
    op.iop_wnew = g->iop_wnew.p;
    op.mark = g->mark_out.p;
    op.mark_gen = g->mark_generation;
    op.touched = g->touched_out.p;
    op.row_a = g->row_a_out.p;
    op.row_b = g->row_b_out.p;
    op.ctl = dc;
    // ---- the in-rows' operations (they do not read the out-rows) ----
    in_pair_params ip{};
    ip.L = L;
    ip.ins_w = ins_w;
    ip.upsert = op.upsert;
    ip.keys = g->keys_b.p + N;
    ip.vals = g->vals_b.p + N;
    ip.N = N;
    ip.vbits = vbits;
    ip.K = K;
    ip.iop_type = g->iop_type.p;
    ip.iop_wold = g->iop_wold.p;
    ip.iop_wnew = g->iop_wnew.p;
    ip.meta = g->in.meta.p;
    ip.base = g->in.base.p;
    ip.dead = g->in.dead.p;
    ip.nbr = g->in.nbr.p;
    ip.w = g->in.w.p;
    ip.slot_cap = g->in.slot_cap;
    ip.pend_alive = g->ipend_alive.p;
    ip.kill = g->ikill.p;
    ip.pend_w = g->ipend_w.p;
    ip.mark = g->mark_in.p;
    ip.mark_gen = g->mark_generation;
    ip.touched = g->touched_in.p;
    ip.row_a = g->row_a_in.p;
    ip.row_b = g->row_b_in.p;
    ip.ctl = dc;
    const unsigned pair_groups = blocks_for(N + 1);
    if (!g->props.allow_parallel) {
      // a simple graph: the in-row pass replays the operations itself, in the same launch
      mark_stage(g, 5, s);
      hipLaunchKernelGGL(k_pairs_simple, dim3(2 * pair_groups), dim3(kBlock), 0, s, op, ip,
                         static_cast<int>(pair_groups));
      DYNG_HIP_TRY(hipGetLastError());
      mark_stage(g, 6, s);
      mark_stage(g, 7, s);
    } else {
      mark_stage(g, 5, s);
      hipLaunchKernelGGL(k_out_pairs, dim3(pair_groups), dim3(kBlock), 0, s, op);
      DYNG_HIP_TRY(hipGetLastError());
      mark_stage(g, 6, s);
      mark_stage(g, 7, s);
      hipLaunchKernelGGL(k_in_pairs, dim3(pair_groups), dim3(kBlock), 0, s, ip);
      DYNG_HIP_TRY(hipGetLastError());
    }
    mark_stage(g, 8, s);
    exclusive_scan4(g->scan_tmp, g->pend_alive.p, g->kill.p, g->ipend_alive.p, g->ikill.p, g->pend_prefix.p,
                    g->kill_prefix.p, g->ipend_prefix.p, g->ikill_prefix.p, N + 1, s);
    mark_stage(g, 9, s);
    // ---- plan the merges: each row's mode, and the staging and tail slots they take ----
    g->merge_lists.reserve(2 * kModes * (N + 1));
    if (wait != nullptr) DYNG_HIP_TRY(hipStreamWaitEvent(s, wait, 0));  // before the rows change
    ap.speculative = N <= kSpeculativeOps;
    if (ap.speculative) {
      merges_speculative(g, N, vbits, s);
    } else {
      merges_exact(g, N, vbits, s, out, false);
    }
  }
  if (wait != nullptr) DYNG_HIP_TRY(hipStreamWaitEvent(s, wait, 0));
  DYNG_HIP_TRY(hipMemcpyAsync(hc, dc, sizeof(apply_control), hipMemcpyDeviceToHost, s));
  // from here on the graph has the batch's vertices (work queued behind reads them)
  g->n = n_after;
  return ap;
}

void apply_finish(dyng_hip_graph* g, const batch_scan& sc, apply_pending& ap, hipStream_t s,
                  dyng_hip_apply_summary* out) {
  const int K = g->K;
  apply_control* hc = g->host_ctl.p;
  apply_control* dc = g->ctl.p;
  const int64_t n_before = ap.n_before;
  const int64_t n_after = g->n;
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (ap.speculative && hc->hold != 0 && hc->error_bits == 0) {
    // the arenas could not take the merges as the device planned them: plan again with the
    // host's compaction and staging growth (nothing was merged; the rows are as the pairing left
    // them)
    hc->hold = 0;
    DYNG_HIP_TRY(hipMemcpyAsync(&dc->hold, &hc->hold, sizeof(int), hipMemcpyHostToDevice, s));
    merges_exact(g, ap.N, ap.vbits, s, out, true);
    if (ap.wait != nullptr) DYNG_HIP_TRY(hipStreamWaitEvent(s, ap.wait, 0));
    DYNG_HIP_TRY(hipMemcpyAsync(hc, dc, sizeof(apply_control), hipMemcpyDeviceToHost, s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    ap.replanned = true;
  }
  if (hc->error_bits != 0) {
    g->n = n_before;
    throw status_error(DYNG_HIP_DEVICE_ERROR, "dyng: the device apply reported an internal error",
                       static_cast<uint32_t>(hc->error_bits));
  }
  g->m += static_cast<int64_t>(hc->inserted_edges) - static_cast<int64_t>(hc->deleted_edges);
  for (int k = 0; k < K; ++k) g->wmin[k] = std::min<int64_t>(g->wmin[k], sc.wmin[k]);
  g->out.tail = hc->tail_out;
  g->in.tail = hc->tail_in;
  g->out.garbage += hc->garbage_out;
  g->in.garbage += hc->garbage_in;
  bool dirty = false;
  for (int k = 0; k < K; ++k) {
    const int64_t target = hc->col_max[k];
    const int64_t before = target == g->wmax[k] ? g->wmax_count[k] : 0;
    g->wsum[k] += hc->col_dsum[k];
    g->wmax[k] = target;
    g->wmax_count[k] = before + hc->col_dcount[k];
    if (g->wmax_count[k] <= 0) dirty = true;
  }
  g->version += 1;
  if (out != nullptr) {
    out->inserted_edges = static_cast<int64_t>(hc->inserted_edges);
    out->updated_edges = static_cast<int64_t>(hc->updated_edges);
    out->deleted_edges = static_cast<int64_t>(hc->deleted_edges);
    out->ignored_deletions = static_cast<int64_t>(hc->ignored_deletions);
    out->dropped_self_loops = static_cast<int64_t>(hc->dropped_self_loops);
    out->ignored_insertions = static_cast<int64_t>(hc->ignored_insertions);
    out->inserted_vertices = n_after - n_before;
    out->num_vertices_after = n_after;
    out->relocated_rows = static_cast<int64_t>(hc->relocated_rows);
    stage_times(g, ap.N > 0 && hc->error_bits == 0, out);
    g->last.summary = *out;
  }
  g->last.valid = true;
  // the last edge of the largest weight left: recompute the column summaries (reuses the control
  // block, so only after the counters above were read)
  if (dirty) recompute_columns(g, s);
}

void apply_scanned(dyng_hip_graph* g, const batch_scan& sc, hipStream_t s, dyng_hip_apply_summary* out,
                   hipEvent_t wait) {
  apply_pending ap = apply_launch(g, sc, s, out, wait);
  apply_finish(g, sc, ap, s, out);
}

const int* apply_hold(const dyng_hip_graph* g) { return &g->ctl.p->hold; }

}  // namespace dyng_hip

// ---------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------

using namespace dyng_hip;

extern "C" int dyng_hip_graph_create_from_edges(int device, const dyng_hip_graph_props* props, int64_t n,
                                                int64_t count, int32_t num_weights, const int32_t* src,
                                                const int32_t* dst, const int32_t* w, int32_t from_device,
                                                void* stream, dyng_hip_graph** out) {
  return guarded([&] {
    if (out == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: null output handle");
    for (int64_t i = 0; !from_device && i < count; ++i) {
      if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n) {
        fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: edge " + std::to_string(i) + " (" + std::to_string(src[i]) + ", " +
                                            std::to_string(dst[i]) + ") is out of range [0, " + std::to_string(n) + ")");
      }
    }
    *out = create_from_edge_list(device, props, n, count, num_weights, src, dst, w, from_device, as_stream(stream));
  });
}

extern "C" int dyng_hip_graph_create(int device, const dyng_hip_graph_props* props, int64_t n, int64_t m,
                                     int32_t num_weights, const int64_t* row_ptr, const int32_t* col,
                                     const int32_t* w, int32_t from_device, void* stream, dyng_hip_graph** out) {
  return guarded([&] {
    if (from_device) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: graph_create takes host rows (use create_from_edges)");
    if (n < 0 || m < 0) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: negative sizes");
    if (row_ptr[0] != 0 || row_ptr[n] != m) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: row offsets do not span the edges");
    std::vector<int32_t> src(static_cast<size_t>(m));
    std::vector<int32_t> wl(static_cast<size_t>(m) * num_weights);
    for (int64_t u = 0; u < n; ++u) {
      if (row_ptr[u + 1] < row_ptr[u]) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: row offsets decrease");
      for (int64_t e = row_ptr[u]; e < row_ptr[u + 1]; ++e) src[e] = static_cast<int32_t>(u);
    }
    for (int64_t e = 0; e < m; ++e) {
      if (col[e] < 0 || col[e] >= n) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: neighbour out of range");
      for (int k = 0; k < num_weights; ++k) wl[e * num_weights + k] = w[k * m + e];
    }
    *out = create_from_edge_list(device, props, n, m, num_weights, src.data(), col, wl.data(), 0, as_stream(stream));
  });
}

extern "C" int dyng_hip_graph_destroy(dyng_hip_graph* g) {
  return guarded([&] { delete g; });
}

extern "C" int dyng_hip_graph_num_vertices(const dyng_hip_graph* g, int64_t* n) {
  return guarded([&] { *n = g->n; });
}

extern "C" int dyng_hip_graph_num_edges(const dyng_hip_graph* g, int64_t* m) {
  return guarded([&] { *m = g->m; });
}

extern "C" int dyng_hip_graph_version(const dyng_hip_graph* g, uint64_t* version) {
  return guarded([&] { *version = g->version; });
}

extern "C" int dyng_hip_graph_column(const dyng_hip_graph* g, int32_t k, int64_t* max_weight, int64_t* weight_sum) {
  return guarded([&] {
    if (k < 0 || k >= g->K) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: objective out of range");
    *max_weight = g->wmax[k];
    *weight_sum = g->wsum[k];
  });
}

extern "C" int dyng_hip_graph_download(const dyng_hip_graph* g, int64_t* row_ptr, int32_t* col, int32_t* w) {
  return guarded([&] {
    scoped_device guard(g->device);
    DYNG_HIP_TRY(hipDeviceSynchronize());
    copy_rows_to_host(g->out, g->n, g->K, row_ptr, col, w);
  });
}

extern "C" int dyng_hip_graph_download_in(const dyng_hip_graph* g, int64_t* row_ptr, int32_t* col, int32_t* w) {
  return guarded([&] {
    scoped_device guard(g->device);
    DYNG_HIP_TRY(hipDeviceSynchronize());
    copy_rows_to_host(g->in, g->n, g->K, row_ptr, col, w);
  });
}

extern "C" int dyng_hip_graph_apply(dyng_hip_graph* g, const dyng_hip_batch* batch, void* stream,
                                    dyng_hip_apply_summary* summary) {
  return guarded([&] {
    if (g == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: null graph");
    scoped_device guard(g->device);
    dyng_hip_apply_summary tmp{};
    batch_scan sc;
    scan_batch(g, batch, as_stream(stream), sc);
    apply_scanned(g, sc, as_stream(stream), &tmp);
    if (summary != nullptr) *summary = tmp;
  });
}

extern "C" int dyng_hip_graph_changes(const dyng_hip_graph* g, dyng_hip_changes* out) {
  return guarded([&] {
    out->del_src = g->del_src.p;
    out->del_dst = g->del_dst.p;
    out->num_del = g->num_del_slots;
    out->ins_src = g->ins_src.p;
    out->ins_dst = g->ins_dst.p;
    out->ins_increased = g->ins_inc.p;
    out->num_ins = g->num_ins_slots;
    out->num_weights = g->K;
  });
}
