// The sssp engine of the HIP path: one persistent kernel per compute() / update() (the fused
// engine of cpp/src/algorithms/sssp/fused.cuh:155-495, re-designed for CDNA4).
//
// Phases, separated by grid barriers and decided on the device:
//   A  pack the old tree into (distance << b | parent) words; roots = heads v of changed edges
//      (u,v) with parent[v] == u (fused.cuh:178-197); roots start the invalidated list.
//   B  the subtrees under the roots (fused.cuh:199-234 computes them by pointer jumping over all
//      n vertices). A tree this engine produced is walked top-down instead along its tree edges
//      (out-edges (u, w) with parent[w] == u), touching only the invalidated region: each
//      workgroup walks its tile of the frontier down several levels on its own (levels in LDS)
//      between grid barriers; a walk deeper than kWalkRounds such rounds hands its unfinished
//      part to pointer jumping. A tree handed in from outside keeps pointer jumping (over
//      compacted lists of the vertices still jumping), which also detects parent cycles.
//   C  invalidate (word INF, stamp), then the insertion heads (fused.cuh:236-256).
//   D  pull: each invalidated vertex takes its best in-neighbour offer (fused.cuh:258-285). For a
//      tree this engine produced, an insertion head's old word was already the best over its
//      old in-edges, so it is offered only the inserted edges into it; a tree from outside pulls
//      over every in-edge of the heads too.
//   E  near-far propagation (fused.cuh:288-417): the relax iteration expands the near frontier
//      with a packed 64-bit atomicMin per edge.
//   F  unpack, writing only changed (distance, parent) pairs and counting them
//      (fused.cuh:419-494); the distance-only mode recovers lowest-id parents.
//
// Rows are expanded edge-balanced: a workgroup takes a tile of the list, scans the row lengths in
// LDS and deals the tile's edge slots out one per thread, so a thread's critical path is one edge
// rather than one row. Rows longer than `light` are cut into chunks that whole waves expand: in the
// walk and the pull after a barrier of their own, in the relax phase at the start of the next
// iteration. Per-iteration list counters live in rings of three slots, so a relax iteration needs
// one barrier. Workgroups exchange data only through agent-scope
// atomics and loads/stores that bypass the non-coherent caches, so the grid barrier itself needs
// no cache write-back or invalidation (the distance-only unpack, which reads plain stores of
// other workgroups, uses the fenced barrier).
#include "internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>

namespace dyng_hip {

namespace {

constexpr int kBlock = 1024;  ///< 16 waves per workgroup; 8 waves/SIMD fit two per CU, one is launched
constexpr int kWaves = kBlock / 64;
constexpr int kWalkGlobal = 2;   ///< first rounds of the subtree walk: one level each
constexpr int kWalkLocal = 16;   ///< later rounds: levels a workgroup walks on its own between grid barriers
constexpr int kWalkRounds = kWalkGlobal + 8;  ///< grid rounds of the walk before pointer jumping takes over
constexpr int kRelaxGlobal = 2;  ///< relax iterations of a repair before local rounds may start
constexpr int kRelaxLocal = 16;  ///< local rounds of a workgroup in one relax iteration
constexpr int kRelaxNarrow = 128; ///< a near frontier of at most this many vertices per workgroup is narrow
/// A workgroup's local levels of one walk round or relax iteration stop once they have expanded
/// this many edges; what they claimed next goes to the next grid round, spread over the grid (a
/// local level that reaches a bushy part of the graph would otherwise keep the grid waiting on one
/// workgroup: measured on RMAT-22, one local relax iteration of 124 us where the others took 12).
constexpr int kLocalEdgeBudget = 8192;
constexpr int kLocalGenerations = 1600000000;  ///< no local relax iteration starts past this generation

struct control {
  int ring_near[3];
  int ring_chunk[3];
  int ring_jump[3];
  int far_count[2];
  int cand_count;
  int heads_count;
  int frontier_count;
  int touched_count;  ///< entries of the unpack's list of vertices first stamped by this run's claims
  int held;           ///< 1: the run did nothing (the apply ahead of it was held back, see engine_params::hold)
  int roots;
  int overflow;
  int invalidated;
  int rounds;
  int still_jumping;
  int iterations;
  int epochs;
  int generation;
  long long pushes;
  u64 min_frontier;
  u64 minimum[2];
  unsigned long long affected;
  unsigned long long relax_edges;
  unsigned long long relax_improved;
  unsigned long long pull_edges;
  unsigned long long jump_items;
  unsigned long long pull_vertices;
  unsigned long long t_phase[8];  ///< s_memrealtime ticks: start, pull, near-far, unpack, after A, B, C, end
  unsigned int error_bits;
};

struct engine_params {
  int32_t n;
  const int2* out_meta;
  const int32_t* out_base;  ///< per vertex: the length of the row's sorted base run
  const int32_t* out_nbr;
  const int32_t* out_w;
  const int2* in_meta;
  const int32_t* in_nbr;
  const int32_t* in_w;
  const int32_t* del_src;
  const int32_t* del_dst;
  int64_t num_del;
  const int32_t* ins_src;
  const int32_t* ins_dst;
  const uint8_t* ins_inc;
  const int32_t* ins_fw;  ///< null, or each insertion's pair weights after the batch (simple graphs)
  int32_t inc_stride;
  int32_t inc_k;
  int64_t num_ins;
  int32_t source;           ///< the source, or -1 with several (sources[0, num_sources))
  const int32_t* sources;
  int64_t num_sources;
  int32_t from_scratch;
  int32_t count_changes;
  int32_t check_cycles;
  packing pk;
  u64 max_distance;
  u64 delta;
  int32_t max_rounds;
  int32_t generation;
  long long* distances;
  int32_t* parents;
  u64* packed;
  int* stamp;
  int* in_far;
  int* flag;
  int* ancestor;
  int32_t resident; ///< update: `packed` already holds the old tree's words (no packing)
  int* cand;
  int* frontier;
  int* near_a;
  int* near_b;
  int* far_a;
  int* far_b;
  int* touched;  ///< a repair's vertices whose words its claims may have changed (see claim_first)
  int4* chunks;
  int64_t chunk_cap;
  control* ctl;
  grid_sync_state* bar;
  int32_t light;
  int32_t chunk_edges;
  unsigned long long barrier_ticks;  ///< longest wait at a grid barrier (s_memrealtime ticks)
  const int* hold;  ///< null, or a flag of the apply queued ahead: nonzero, the run does nothing
};

/// Append v to a list when pred (wave-aggregated reservation).
DYNG_DEV void append(int* list, int* counter, bool pred, int32_t v) {
  const int slot = wave_append(counter, pred);
  if (pred) st(&list[slot], v);
}

constexpr int kStage = 2048;  ///< entries a workgroup stages per list between flushes

/// Appends staged in LDS. A list's counter is one address that every workgroup hits; device-scope
/// atomics on one address serialize (about 10 ns each on MI355X), so one reservation per wave
/// costs microseconds per phase on a big frontier and milliseconds on a pass over every vertex.
/// Waves reserve LDS slots instead (an LDS atomic), and stage_flush reserves the global range of
/// the workgroup's staged entries with one atomicAdd per list. Entries past the stage's capacity
/// go to the list directly.
struct stage_shared {
  int count[3];
  int base[3];
  int item[3][kStage];
  int* list2;     ///< list 2 is always the unpack's list (engine_params::touched)
  int* counter2;  ///< and its counter
};

/// Stage v for list k (`list`, `counter`) when pred. Every active lane of the wave calls it.
DYNG_DEV void stage_append(stage_shared& S, int k, bool pred, int32_t v, int* list, int* counter) {
  const u64 mask = __ballot(pred);
  if (mask == 0) return;  // uniform per wave
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  int b = 0;
  if (lane() == leader) b = atomicAdd(&S.count[k], __popcll(mask));
  b = __shfl(b, leader);
  const int slot = b + __popcll(mask & lanes_below());
  if (pred && slot < kStage) S.item[k][slot] = v;
  append(list, counter, pred && slot >= kStage, v);
}

/// Publish the staged entries of list 0, of list 1 when list1 is not null, and of list 2 (the
/// unpack's list). Every thread of the workgroup calls it; the stage is empty afterwards.
DYNG_DEV void stage_flush(stage_shared& S, int* list0, int* counter0, int* list1 = nullptr, int* counter1 = nullptr) {
  __syncthreads();  // every staged entry is in LDS
  if (threadIdx.x < 3) {
    const int k = threadIdx.x;
    int* counter = k == 0 ? counter0 : k == 1 ? counter1 : S.counter2;
    const int cnt = counter != nullptr ? min(S.count[k], kStage) : 0;
    S.count[k] = cnt;
    S.base[k] = cnt > 0 ? atomicAdd(counter, cnt) : 0;
  }
  __syncthreads();
  for (int k = 0; k < 3; ++k) {
    int* list = k == 0 ? list0 : k == 1 ? list1 : S.list2;
    const int cnt = S.count[k], b = S.base[k];
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) st(&list[b + i], S.item[k][i]);
  }
  __syncthreads();  // every thread has read count and base
  if (threadIdx.x < 3) S.count[threadIdx.x] = 0;
  __syncthreads();  // the reset lands before the next stage_append
}

/// Claim v for the list of generation `generation` (the stamp exchange): true when this call
/// claimed it. `first`: v carried no stamp of this run before (stamps of this run are >= its first
/// generation p.generation), so the unpack's list must take it.
DYNG_DEV bool claim_first(const engine_params& p, int32_t v, int generation, bool& first) {
  const int old = atomicExch(&p.stamp[v], generation);
  first = old < p.generation;
  return old != generation;
}

/// A chunk {vertex, begin, end} of the chunk list, written and read with agent-scope operations.
DYNG_DEV void st_chunk(int4* ch, int32_t v, int32_t b, int32_t e) {
  u64* q = reinterpret_cast<u64*>(ch);
  st(&q[0], static_cast<u64>(static_cast<uint32_t>(v)) | (static_cast<u64>(static_cast<uint32_t>(b)) << 32));
  st(&q[1], static_cast<u64>(static_cast<uint32_t>(e)));
}
DYNG_DEV int4 ld_chunk(const int4* ch) {
  const u64* q = reinterpret_cast<const u64*>(ch);
  const u64 a = ld(&q[0]), b = ld(&q[1]);
  return make_int4(static_cast<int32_t>(static_cast<uint32_t>(a)), static_cast<int32_t>(static_cast<uint32_t>(a >> 32)),
                   static_cast<int32_t>(static_cast<uint32_t>(b)), 0);
}


/// Four stamps at once, as two agent-scope 8-byte loads.
DYNG_DEV int4 ld_int4(const int4* q4) {
  const u64* q = reinterpret_cast<const u64*>(q4);
  const u64 a = ld(&q[0]), b = ld(&q[1]);
  return make_int4(static_cast<int32_t>(static_cast<uint32_t>(a)), static_cast<int32_t>(static_cast<uint32_t>(a >> 32)),
                   static_cast<int32_t>(static_cast<uint32_t>(b)), static_cast<int32_t>(static_cast<uint32_t>(b >> 32)));
}

/// Cut the row [begin, begin + len) of vertex v into chunks of `chunk` edges (one wave each).
DYNG_DEV void emit_chunks_to(const engine_params& p, int4* out, int64_t cap, int* counter, int32_t v, int32_t begin,
                              int32_t len) {
  const int pieces = (len + p.chunk_edges - 1) / p.chunk_edges;
  const int base = atomicAdd(counter, pieces);
  for (int c = 0; c < pieces; ++c) {
    if (base + c >= cap) {
      atomicOr(&p.ctl->error_bits, DYNG_HIP_DEVICE_ERROR_CAPACITY);
      return;
    }
    const int b = begin + c * p.chunk_edges;
    const int e = min(begin + len, b + p.chunk_edges);
    st_chunk(&out[base + c], v, b, e);
  }
}
DYNG_DEV void emit_chunks(const engine_params& p, int* counter, int32_t v, int32_t begin, int32_t len) {
  emit_chunks_to(p, p.chunks, p.chunk_cap, counter, v, begin, len);
}

/// The workgroup's view of one tile of a vertex list: per item the first edge slot of its rows in
/// the tile (exclusive prefix of the lengths), the row offset, the vertex and a 64-bit word.
struct tile_shared {
  int start[kBlock + 1];
  int off[kBlock];
  int vert[kBlock];
  u64 word[kBlock];
  int scan[kWaves + 1];
};

/// Exclusive prefix sum of x over the workgroup; `total` gets the sum. Every thread calls it.
DYNG_DEV int block_scan(tile_shared& L, int x, int& total) {
  int v = x;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(v, off);
    if (lane() >= off) v += y;
  }
  const int w = static_cast<int>(threadIdx.x) >> 6;
  if (lane() == 63) L.scan[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int i = 0; i < kWaves; ++i) {
      const int t = L.scan[i];
      L.scan[i] = run;
      run += t;
    }
    L.scan[kWaves] = run;
  }
  __syncthreads();
  total = L.scan[kWaves];
  return L.scan[w] + v - x;
}

/// Edge slots a lane takes per step of an expansion: their loads are issued together (each slot's
/// row entry, then the words it points at), so a lane waits for one round of gathers per K slots
/// instead of one per slot.
constexpr int kSlots = 1;

/// Edge-balanced expansion of one tile of `cnt` items (cnt <= kBlock): `load(i, v, off, len,
/// word)` fills thread i's item (i < cnt; len 0: nothing to expand, e.g. a long row sent to the
/// chunk list), and the edge slots of the tile's rows are dealt out to the workgroup's threads,
/// kSlots slots per thread per step (slot j of a step kBlock apart from slot j - 1, so each j is
/// one coalesced run per wave): `edges(act, item, slot)` runs with every lane of a wave, on arrays
/// of kSlots (act false on padding). `done(valid, item)` runs on every thread after the tile's
/// edges. Every thread of the workgroup calls it.
template <class Load, class Edges, class Done>
DYNG_DEV void expand_tile(tile_shared& L, int cnt, Load load, Edges edges, Done done) {
  const int i = static_cast<int>(threadIdx.x);
  int32_t v = -1, off = 0, len = 0;
  u64 word = packed_inf;
  if (i < cnt) load(i, v, off, len, word);
  int total = 0;
  const int start = block_scan(L, len, total);
  if (i < cnt) {
    L.start[i] = start;
    L.off[i] = off;
    L.vert[i] = v;
    L.word[i] = word;
  }
  if (i == 0) L.start[cnt] = total;
  __syncthreads();
  for (int base = i - lane(); base < total; base += kSlots * kBlock) {
    bool act[kSlots];
    int item[kSlots], slot[kSlots];
#pragma unroll
    for (int j = 0; j < kSlots; ++j) {
      const int e = base + j * kBlock + lane();
      act[j] = e < total;
      int lo = 0;
      if (act[j]) {
        int hi = cnt;  // the last item whose first slot is <= e (its row holds e)
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (L.start[mid] <= e) {
            lo = mid;
          } else {
            hi = mid;
          }
        }
      }
      item[j] = lo;
      slot[j] = act[j] ? L.off[lo] + (e - L.start[lo]) : 0;
    }
    edges(act, item, slot);
  }
  __syncthreads();
  done(i < cnt, i);
  __syncthreads();
}

/// Edge-balanced expansion of list items [0, count): the items are cut into one tile per
/// workgroup (up to kBlock items), each expanded by expand_tile with `load(index, ...)` given the
/// item's index in the list, and `flush()` runs on every thread of the workgroup after each tile
/// (to publish staged appends). Uniform per workgroup: every thread calls it.
template <class Load, class Edges, class Done, class Flush>
DYNG_DEV void tile_rows(tile_shared& L, int count, Load load, Edges edges, Done done, Flush flush) {
  const int G = static_cast<int>(gridDim.x);
  const int ts = min(kBlock, max(1, (count + G - 1) / G));
  for (long long t0 = static_cast<long long>(blockIdx.x) * ts; t0 < count; t0 += static_cast<long long>(G) * ts) {
    const int cnt = static_cast<int>(min(static_cast<long long>(ts), count - t0));
    const int first = static_cast<int>(t0);
    expand_tile(
        L, cnt, [&](int i, int32_t& v, int32_t& off, int32_t& len, u64& word) { load(first + i, v, off, len, word); },
        edges, done);
    flush();
  }
}

/// w's word changed in this run without its distance falling (a lower parent id at the same
/// distance). Every vertex whose word a repair changes ends up with a stamp of this run (stamps
/// of this run are >= its first generation p.generation): an invalidated vertex or one whose
/// distance fell was claimed for a list; this one gets p.generation itself, which is below every
/// generation a claim compares with, so it claims nothing. The unpack visits the stamped ones.
DYNG_DEV bool mark_changed(const engine_params& p, int32_t w) { return atomicMax(&p.stamp[w], p.generation) < p.generation; }

/// One relaxation of edge (u, w) of weight wt with u at distance du (fused.cuh:341-362), given
/// w's word as read before (`seen`; it only falls during the relax, so a stale read is never
/// below the truth and only costs an atomic): the candidate word joins the next near list or the
/// far pile when it lowered w's word. Returns the lists to append w to.
DYNG_DEV void relax_decide(const engine_params& p, bool active, int32_t u, u64 du, int32_t w, int32_t wt, u64 seen,
                           u64 threshold, int generation, unsigned long long* improved, bool& to_near,
                           bool& to_far, bool& first) {
  to_near = false;
  to_far = false;
  first = false;
  if (active) {
    const u64 nd = du + static_cast<u64>(wt);
    if (nd <= p.max_distance) {
      const u64 cand = p.pk.pack(nd, u);
      if (cand < seen) {
        const u64 old = atomicMin(&p.packed[w], cand);
        // Any lowered word joins a list, also one lowered only to a lower parent id at the same
        // distance (fused.cuh:341-362 lists only a lower distance): the extra expansion is rare
        // and offers nothing new, and being claimed stamps the vertex for the unpack.
        if (cand < old) {
          atomicAdd(improved, 1ULL);
          // both flags as values (a store through a reference chosen by the branch would keep
          // them in scratch memory); a far vertex is stamped when it leaves the far pile
          const bool near = nd < threshold;
          to_near = near && claim_first(p, w, generation, first);
          to_far = !near && atomicExch(&p.in_far[w], 1) == 0;
        }
      }
    }
  }
}

/// The subtree walk's and the local relax rounds' levels in LDS: two levels of up to kBlock
/// vertices.
struct walk_shared {
  int count[2];
  int item[2][kBlock];
};

/// Where a relax step sends the vertices it claimed for the near frontier. In a global iteration
/// (`local` null), the next near list. In a workgroup's local rounds, its next local level first
/// (`local`->item[q]), and the next near list only when that is full; a vertex enters that list
/// once per iteration there (`listed[v]` takes the iteration's `tag` when it does), since the
/// local rounds' claims use a generation per workgroup and round and may claim it again.
struct near_sink {
  walk_shared* local;
  int q;
  int* listed;
  int tag;
};

/// kSlots relaxations per lane: the row entries first, then the words they point at, then the
/// atomics and the list appends. Every lane of the wave calls it.
DYNG_DEV void relax_slots(const engine_params& p, const bool (&act)[kSlots], const int32_t (&u)[kSlots],
                          const u64 (&du)[kSlots], const int (&slot)[kSlots], u64 threshold, int generation,
                          stage_shared& S, int* next, int* next_count, int* far, int* far_count,
                          unsigned long long* improved, const near_sink& sink) {
  int32_t w[kSlots], wt[kSlots];
  u64 seen[kSlots];
  bool live[kSlots];
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    w[j] = act[j] ? p.out_nbr[slot[j]] : -1;
    wt[j] = act[j] ? p.out_w[slot[j]] : 0;
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    live[j] = act[j] && w[j] >= 0 && w[j] != p.source;  // a dead entry is negative
    seen[j] = live[j] ? ld(&p.packed[w[j]]) : 0;
  }
#pragma unroll
  for (int j = 0; j < kSlots; ++j) {
    bool to_near, to_far, first;
    relax_decide(p, live[j], u[j], du[j], w[j], wt[j], seen[j], threshold, generation, improved, to_near, to_far,
                 first);
    bool to_list = to_near;
    if (to_near && sink.local != nullptr) {
      const int at = atomicAdd(&sink.local->count[sink.q], 1);
      if (at < kBlock) {
        sink.local->item[sink.q][at] = w[j];
        to_list = false;
      }
    }
    if (to_list && sink.listed != nullptr) to_list = atomicExch(&sink.listed[w[j]], sink.tag) != sink.tag;
    stage_append(S, 0, to_list, w[j], next, next_count);
    stage_append(S, 1, to_far, w[j], far, far_count);
    stage_append(S, 2, first, w[j], S.list2, S.counter2);
  }
}

/// Chunks of long rows, one wave per chunk: `edges(act, u, slot)` with every lane of the wave, on
/// arrays of kSlots slots (64 apart).
template <class Edges>
DYNG_DEV void chunk_rows(const engine_params& p, int nchunks, Edges edges) {
  const int wave = static_cast<int>((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const int waves = static_cast<int>((gridDim.x * blockDim.x) / 64);
  for (int c = wave; c < nchunks; c += waves) {
    const int4 ch = ld_chunk(&p.chunks[c]);
    for (int b = ch.y; b < ch.z; b += 64 * kSlots) {
      bool act[kSlots];
      int32_t u[kSlots];
      int slot[kSlots];
#pragma unroll
      for (int j = 0; j < kSlots; ++j) {
        const int e = b + j * 64 + lane();
        act[j] = e < ch.z;
        u[j] = ch.x;
        slot[j] = act[j] ? e : 0;
      }
      edges(act, u, slot);
    }
  }
}

/// Pull offer of in-edge (u -> v, wt) given u's word: the candidate word, or INF (also for a dead
/// entry, u < 0).
DYNG_DEV u64 offer_of(const engine_params& p, int32_t u, int32_t wt, u64 word) {
  if (u < 0 || word == packed_inf) return packed_inf;
  const u64 nd = p.pk.distance(word) + static_cast<u64>(wt);
  return nd <= p.max_distance ? p.pk.pack(nd, u) : packed_inf;
}

/// Pull offer of in-edge (u -> v, wt): the candidate word, or INF (also for a dead entry, u < 0).
DYNG_DEV u64 pull_offer(const engine_params& p, int32_t u, int32_t wt) {
  return u < 0 ? packed_inf : offer_of(p, u, wt, ld(&p.packed[u]));
}

/// The best offer to v over the alive entries (u, v) of u's out-row (both sorted runs), or INF.
DYNG_DEV u64 edge_offer(const engine_params& p, int32_t u, int32_t v) {
  const u64 word = ld(&p.packed[u]);
  if (word == packed_inf) return packed_inf;
  const u64 du = p.pk.distance(word);
  const int2 mt = p.out_meta[u];
  const int bl = p.out_base[u];
  const int32_t* row = p.out_nbr + mt.x;
  u64 best = packed_inf;
  const int runs[3] = {0, bl, mt.y};
  for (int r = 0; r < 2; ++r) {
    // the first entry of the run whose neighbour (dead or alive) is >= v
    const int lo = row_lower_bound(row, runs[r], runs[r + 1], v);
    for (int x = lo; x < runs[r + 1] && (row[x] & 0x7fffffff) == v; ++x) {
      if (row[x] < 0) continue;  // dead
      const u64 nd = du + static_cast<u64>(p.out_w[mt.x + x]);
      if (nd <= p.max_distance) {
        const u64 o = p.pk.pack(nd, u);
        best = o < best ? o : best;
      }
    }
  }
  return best;
}

/// The offer of insertion i's pair (u, v) on a simple graph: u's distance plus the pair's weight
/// after the batch, which the apply recorded (the pair is one entry of u's row, so no search).
DYNG_DEV u64 final_offer(const engine_params& p, int32_t u, int64_t i) {
  const u64 word = ld(&p.packed[u]);
  if (word == packed_inf) return packed_inf;
  const int32_t w = p.ins_fw[i * p.inc_stride + p.inc_k];
  if (w < 1) return packed_inf;  // the batch left no (u, v)
  const u64 nd = p.pk.distance(word) + static_cast<u64>(w);
  return nd <= p.max_distance ? p.pk.pack(nd, u) : packed_inf;
}

/// Lower v's word to its best pull offer; v joins the frontier if its distance fell. The
/// smallest lowered distance goes to *minimum (the first near-far threshold).
DYNG_DEV void pull_commit(const engine_params& p, bool pred, int32_t v, u64 best, int generation, stage_shared& S,
                          int* counter, u64* minimum) {
  bool add = false, first = false;
  u64 low = packed_inf;
  if (pred) {
    const u64 current = ld(&p.packed[v]);
    if (best < current) {
      const u64 old = atomicMin(&p.packed[v], best);
      if (p.pk.distance(best) < p.pk.distance(old)) {
        low = p.pk.distance(best);
        add = claim_first(p, v, generation, first);
      } else if (best < old) {
        first = mark_changed(p, v);
      }
    }
  }
  stage_append(S, 0, add, v, p.frontier, counter);
  stage_append(S, 2, first, v, S.list2, S.counter2);
  low = wave_min(low);
  if (lane() == 0 && low != packed_inf) atomicMin(minimum, low);
}

}  // namespace

/// Leave the kernel when the grid barrier gave up (see grid_barrier); uniform per workgroup.
/// DYNG_GRID_SYNC separates phases whose shared data is only touched with agent-scope
/// operations; DYNG_GRID_SYNC_FENCED also orders plain loads and stores.
#define DYNG_GRID_SYNC_AS(fenced)                                                   \
  do {                                                                              \
    if (!grid_barrier<fenced>(p.bar, p.barrier_ticks)) {                            \
      if (threadIdx.x == 0) atomicOr(&c->error_bits, DYNG_HIP_DEVICE_ERROR_TIMEOUT); \
      return;                                                                       \
    }                                                                               \
  } while (0)
#define DYNG_GRID_SYNC() DYNG_GRID_SYNC_AS(false)
#define DYNG_GRID_SYNC_FENCED() DYNG_GRID_SYNC_AS(true)

/// Clear the grid barrier (a kernel, not a memset, so the launch can live in a hipGraph).
/// Before each run (one workgroup): the barrier state and the control block zeroed, the control
/// block's minima at INF (a kernel rather than a host copy: it costs the stream no blit).
static_assert(sizeof(control) % 8 == 0 && sizeof(grid_sync_state) % 8 == 0, "reset in 8-byte words");
__global__ void k_reset_barrier(grid_sync_state* bar, control* ctl) {
  unsigned long long* w = reinterpret_cast<unsigned long long*>(bar);
  for (unsigned int i = threadIdx.x; i < sizeof(grid_sync_state) / 8; i += blockDim.x) w[i] = 0;
  unsigned long long* c = reinterpret_cast<unsigned long long*>(ctl);
  for (unsigned int i = threadIdx.x; i < sizeof(control) / 8; i += blockDim.x) c[i] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    ctl->min_frontier = packed_inf;
    ctl->minimum[0] = packed_inf;
    ctl->minimum[1] = packed_inf;
  }
}

/// Two builds of the persistent kernel: kWaves = 8 waves per SIMD keeps it at 64 VGPRs, so two
/// 1024-thread workgroups fit a CU (solves from the sources), at the price of 17 VGPRs spilled to
/// scratch; kWaves = 4 gives it 128 (it takes 85, nothing spilled) for grids of one workgroup per
/// CU (repairs).
template <int kWaves>
__global__ void __launch_bounds__(kBlock, kWaves) sssp_fused_kernel(engine_params p) {
  // queued right behind an apply whose merges were held back: nothing to do (uniform)
  if (p.hold != nullptr && *p.hold != 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) p.ctl->held = 1;
    return;
  }
  __shared__ tile_shared L;
  // work counters of the workgroup, in LDS: relax edges, improving atomics, pull edges, affected,
  // expansions of local relax rounds
  __shared__ unsigned long long counts[5];
  __shared__ stage_shared S;
  __shared__ walk_shared W;
  if (threadIdx.x < 5) counts[threadIdx.x] = 0;
  if (threadIdx.x < 3) S.count[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    S.list2 = p.touched;
    S.counter2 = &p.ctl->touched_count;
  }
  __syncthreads();
  unsigned long long* const relax_edges = &counts[0];
  unsigned long long* const relax_improved = &counts[1];
  unsigned long long* const pull_edges = &counts[2];
  unsigned long long* const affected = &counts[3];
  unsigned long long* const local_pushes = &counts[4];
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int threads = gridDim.x * blockDim.x;
  const int n = p.n;
  control* c = p.ctl;
  int generation = p.generation;
  int frontier_count = 0;
  int step = 0, cstep = 0;  // ring positions (uniform across the grid)
  unsigned long long jump_items = 0;
  if (tid == 0) c->t_phase[0] = __builtin_amdgcn_s_memrealtime();

  if (p.from_scratch) {
    if (p.source >= 0) {
      for (int v = tid; v < n; v += threads) st(&p.packed[v], v == p.source ? p.pk.pack(0, -1) : packed_inf);
      if (tid == 0) st(&p.frontier[0], p.source);
      frontier_count = 1;
    } else {
      // several sources (each at distance 0 with no parent; weights are >= 1, so no offer ever
      // beats a source's word and no source needs skipping by id)
      for (int v = tid; v < n; v += threads) st(&p.packed[v], packed_inf);
      DYNG_GRID_SYNC();
      for (int64_t i = tid; i < p.num_sources; i += threads) {
        st(&p.packed[p.sources[i]], p.pk.pack(0, -1));
        st(&p.frontier[i], p.sources[i]);
      }
      frontier_count = static_cast<int>(p.num_sources);
    }
    DYNG_GRID_SYNC();
  } else {
    // ---- A: pack the old tree (unless the workspace still holds its words); ancestors (pointer
    // jumping only); roots ----
    for (int v = tid; !p.resident && v < n; v += threads) {
      const long long d = p.distances[v];
      const int32_t par = p.parents[v];
      u64 word = packed_inf;
      if (d < sssp_infinity / 2) {
        if (d < 0 || static_cast<u64>(d) > p.max_distance) {
          st(&c->overflow, 1);
        } else {
          word = p.pk.pack(static_cast<u64>(d), par);
        }
      }
      st(&p.packed[v], word);
      if (p.check_cycles) st(&p.ancestor[v], par);
    }
    for (int64_t i = tid; i < p.num_del + p.num_ins; i += threads) {
      int32_t u, v;
      bool changed;
      if (i < p.num_del) {
        u = p.del_src[i];
        v = p.del_dst[i];
        changed = v >= 0;
      } else {
        const int64_t j = i - p.num_del;
        u = p.ins_src[j];
        v = p.ins_dst[j];
        changed = v >= 0 && p.ins_inc[j * p.inc_stride + p.inc_k] != 0;
      }
      const bool root = changed && p.parents[v] == u && atomicExch(&p.flag[v], 1) == 0;
      if (root) st(&p.stamp[v], generation + 1);
      append(p.cand, &c->cand_count, root, v);
    }
    DYNG_GRID_SYNC();
    if (tid == 0) c->t_phase[4] = __builtin_amdgcn_s_memrealtime();
    if (ld(&c->overflow) != 0) return;  // uniform: the host reports invalid_input
    const int roots = uniform(ld(&c->cand_count));
    const bool work = p.num_del + p.num_ins > 0;
    int inv = roots;  // the invalidated vertices: cand[0, inv)
    bool jump = work && p.check_cycles;
    int jump_slot = 0;  // the first ring slot pointer jumping counts in (a slot already zero)
    if (work && !p.check_cycles && roots > 0) {
      // ---- B (tree from this engine): the subtrees under the roots, down the tree edges
      // (parent[w] == u for an out-edge (u, w)); every tree edge of the old tree is an edge of
      // the graph unless this batch deleted it, and then its head is a root. A tree edge is
      // claimed exactly once (the child's flag), so no workgroup needs another's level: each takes
      // a tile of the frontier and walks it down on its own, its levels in LDS, a level's rows
      // edge-balanced over its threads: one level in each of the first kWalkGlobal rounds (the
      // widest levels, next to the roots, stay spread over the grid: measured on RMAT-22 and
      // ER-64K, two such rounds beat one and eight), then up to kWalkLocal levels per round (the
      // narrow levels of a deep tree: the 4096^2 grid's walk 2.0 -> 0.4 ms). What a workgroup did
      // not reach (a level past the cap or past the LDS queue, and the children of rows long
      // enough for the chunk list) is the next round's frontier, after a grid barrier. Every flagged vertex joins the
      // candidate list after the roots. A walk still going after kWalkRounds rounds (a deep tree)
      // is finished by pointer jumping ----
      // the frontier: a range [lo, hi) of the candidate list while rounds walk one level (the
      // level's children are appended right after it), then a list of its own
      const int* front = p.cand;
      int front_count = roots;
      int lo = 0, hi = roots;
      int rounds = 0;
      // kSlots out-edges (u, w) per lane: the entries, then parent[w] (the tree is read-only
      // here), then the flag claims of the tree children (rare: one in-edge of each child)
      auto claim_children = [&](const bool (&act)[kSlots], const int32_t (&u)[kSlots], const int (&slot)[kSlots],
                                int32_t (&w)[kSlots], bool (&add)[kSlots]) {
        int32_t par[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) w[j] = act[j] ? p.out_nbr[slot[j]] : -1;  // dead entries are negative
#pragma unroll
        for (int j = 0; j < kSlots; ++j) par[j] = w[j] >= 0 ? p.parents[w[j]] : -2;
#pragma unroll
        for (int j = 0; j < kSlots; ++j) {
          add[j] = w[j] >= 0 && par[j] == u[j] && atomicExch(&p.flag[w[j]], 1) == 0;
          if (add[j]) st(&p.stamp[w[j]], generation + 1);
        }
      };
      auto row_of = [&](int32_t v, int32_t& off, int32_t& len, int* chunk_count) {
        const int2 mt = p.out_meta[v];
        if (mt.y <= p.light) {
          off = mt.x;
          len = mt.y;
        } else {
          emit_chunks(p, chunk_count, v, mt.x, mt.y);
        }
      };
      while (front_count > 0 && rounds < kWalkRounds) {
        // the first rounds walk one level each (a bushy tree's levels spread over the grid); the
        // rest kWalkLocal levels each (a deep tree's narrow levels walked without barriers)
        const bool local = rounds >= kWalkGlobal;
        int* spill = (rounds & 1) ? p.near_b : p.near_a;
        int* spill_count = &c->ring_jump[rounds % 3];
        int* chunk_count = &c->ring_chunk[cstep % 3];
        if (tid == 0) {
          st(&c->ring_jump[(rounds + 1) % 3], 0);
          st(&c->ring_chunk[(cstep + 1) % 3], 0);
        }
        ++rounds;
        if (!local) {
          tile_rows(
              L, front_count,
              [&](int idx, int32_t& v, int32_t& off, int32_t& len, u64&) {
                v = ld(&front[idx]);
                row_of(v, off, len, chunk_count);
              },
              [&](const bool (&act)[kSlots], const int (&item)[kSlots], const int (&slot)[kSlots]) {
                int32_t u[kSlots], w[kSlots];
                bool add[kSlots];
#pragma unroll
                for (int j = 0; j < kSlots; ++j) u[j] = act[j] ? L.vert[item[j]] : -1;
                claim_children(act, u, slot, w, add);
#pragma unroll
                for (int j = 0; j < kSlots; ++j) stage_append(S, 0, add[j], w[j], p.cand, &c->cand_count);
              },
              [&](bool, int) {}, [&] { stage_flush(S, p.cand, &c->cand_count); });
        } else {
          const int G = static_cast<int>(gridDim.x);
          const int ts = min(kBlock, max(1, (front_count + G - 1) / G));
          for (long long t0 = static_cast<long long>(blockIdx.x) * ts; t0 < front_count;
               t0 += static_cast<long long>(G) * ts) {
            const int cnt = static_cast<int>(min(static_cast<long long>(ts), front_count - t0));
            if (static_cast<int>(threadIdx.x) < cnt) W.item[0][threadIdx.x] = ld(&front[t0 + threadIdx.x]);
            if (threadIdx.x == 0) {
              W.count[0] = cnt;
              W.count[1] = 0;
            }
            __syncthreads();
            int q = 0;
            int spent = 0;  // edges this tile's local levels expanded
            for (int level = 0; level < kWalkLocal && spent < kLocalEdgeBudget; ++level) {
              const int items = W.count[q];  // uniform: LDS after a workgroup barrier
              if (items == 0) break;
              expand_tile(
                  L, items,
                  [&](int i, int32_t& v, int32_t& off, int32_t& len, u64&) {
                    v = W.item[q][i];
                    row_of(v, off, len, chunk_count);
                  },
                  [&](const bool (&act)[kSlots], const int (&item)[kSlots], const int (&slot)[kSlots]) {
                    int32_t u[kSlots], w[kSlots];
                    bool add[kSlots];
#pragma unroll
                    for (int j = 0; j < kSlots; ++j) u[j] = act[j] ? L.vert[item[j]] : -1;
                    claim_children(act, u, slot, w, add);
#pragma unroll
                    for (int j = 0; j < kSlots; ++j) {
                      // the child joins the next local level, or the next round's frontier when
                      // the level is full
                      const int at = add[j] ? atomicAdd(&W.count[q ^ 1], 1) : 0;
                      if (add[j] && at < kBlock) W.item[q ^ 1][at] = w[j];
                      stage_append(S, 0, add[j], w[j], p.cand, &c->cand_count);
                      stage_append(S, 1, add[j] && at >= kBlock, w[j], spill, spill_count);
                    }
                  },
                  [&](bool, int) {});
              spent += L.start[items];  // the level's edges (LDS after the expansion's barrier: uniform)
              if (threadIdx.x == 0) {
                W.count[q] = 0;
                W.count[q ^ 1] = min(W.count[q ^ 1], kBlock);
              }
              if (S.count[0] > kStage / 2 || S.count[1] > kStage / 2) {
                stage_flush(S, p.cand, &c->cand_count, spill, spill_count);  // uniform: LDS after a barrier
              }
              __syncthreads();
              q ^= 1;
            }
            // a level the cap left: the next round's frontier
            const int left = W.count[q];
            for (int base = static_cast<int>(threadIdx.x) - lane(); base < left; base += kBlock) {
              const int i = base + lane();
              stage_append(S, 1, i < left, i < left ? W.item[q][i] : -1, spill, spill_count);
            }
            stage_flush(S, p.cand, &c->cand_count, spill, spill_count);
          }
        }
        DYNG_GRID_SYNC();
        // long rows: their chunks, by every workgroup; the children join the next frontier
        const int nchunks = uniform(static_cast<int>(min(static_cast<int64_t>(ld(chunk_count)), p.chunk_cap)));
        ++cstep;
        if (nchunks > 0) {
          chunk_rows(p, nchunks, [&](const bool (&act)[kSlots], const int32_t (&u)[kSlots], const int (&slot)[kSlots]) {
            int32_t w[kSlots];
            bool add[kSlots];
            claim_children(act, u, slot, w, add);
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
              stage_append(S, 0, add[j], w[j], p.cand, &c->cand_count);
              stage_append(S, 1, local && add[j], w[j], spill, spill_count);
            }
          });
          stage_flush(S, p.cand, &c->cand_count, spill, spill_count);
          DYNG_GRID_SYNC();
        }
        if (!local) {
          // the next round starts from this level's children, right after it in the list
          lo = hi;
          hi = uniform(ld(&c->cand_count));
          front = p.cand + lo;
          front_count = hi - lo;
        } else {
          front = spill;
          front_count = uniform(ld(spill_count));
        }
      }
      if (tid == 0) st(&c->rounds, rounds);
      hi = uniform(ld(&c->cand_count));  // every flagged vertex (no append after the barrier)
      jump_items = hi;
      if (front_count > 0) {
        // deep: pointer jumping from every vertex flagged so far (they are cand[0, hi)). Its
        // rounds continue the ring where the walk stopped: slot `rounds % 3` was cleared during
        // the last round and nobody reads it, while the last round's count may still be unread
        // by slower workgroups and must not be touched here.
        for (int v = tid; v < n; v += threads) st(&p.ancestor[v], p.parents[v]);
        DYNG_GRID_SYNC();
        jump = true;
        jump_slot = rounds;
      } else {
        inv = hi;
      }
    }
    if (jump) {
      // ---- B (tree from outside, or a deep walk): pointer jumping over compacted lists of
      // the vertices still jumping; a list that is not empty after ceil(log2 n) + 1 rounds is a
      // parent cycle ----
      int count = n;
      const int* list = nullptr;
      // (selects, not a two-entry array: a dynamically indexed private array lives in scratch)
      for (int r = 0; r < p.max_rounds; ++r) {
        jump_items += count;
        int* out = (r & 1) ? p.near_b : p.near_a;
        int* out_count = &c->ring_jump[(r + jump_slot) % 3];
        if (tid == 0) {
          st(&c->ring_jump[(r + jump_slot + 1) % 3], 0);
          st(&c->rounds, r + 1);
        }
        // whole workgroups per step (t0 is uniform per workgroup), so each step can publish its
        // staged appends
        for (long long t0 = static_cast<long long>(blockIdx.x) * kBlock; t0 < count; t0 += threads) {
          const int i = static_cast<int>(t0) + static_cast<int>(threadIdx.x);
          bool keep = false, flagged = false;
          int32_t v = -1;
          if (i < count) {
            v = list ? ld(&list[i]) : i;
            const int32_t a = ld(&p.ancestor[v]);
            if (a >= 0 && ld(&p.flag[v]) == 0) {
              if (ld(&p.flag[a]) != 0) {
                st(&p.flag[v], 1);
                st(&p.stamp[v], generation + 1);
                flagged = true;
              } else {
                const int32_t next = ld(&p.ancestor[a]);
                st(&p.ancestor[v], next);
                keep = next >= 0;
              }
            }
          }
          stage_append(S, 0, flagged, v, p.cand, &c->cand_count);
          stage_append(S, 1, keep, v, out, out_count);
          stage_flush(S, p.cand, &c->cand_count, out, out_count);
        }
        DYNG_GRID_SYNC();
        count = uniform(ld(out_count));
        list = out;
        if (count == 0) break;
      }
      if (count > 0) {
        if (tid == 0) atomicOr(&c->error_bits, DYNG_HIP_DEVICE_ERROR_PARENT_CYCLE);
        return;  // uniform
      }
      inv = uniform(ld(&c->cand_count));  // nothing appends to it after the last round's barrier
    }
    // ---- C: invalidate the flagged vertices (stamped when flagged) and claim the insertion
    // heads that are not invalidated, in one phase ----
    if (tid == 0) c->t_phase[5] = __builtin_amdgcn_s_memrealtime();
    ++generation;
    for (int i = tid; i < inv; i += threads) {
      const int32_t v = ld(&p.cand[i]);
      st(&p.flag[v], 0);
      st(&p.packed[v], packed_inf);
    }
    if (tid == 0) st(&c->invalidated, inv);
    // A tree from outside: the insertion heads that are not invalidated pull over all their
    // in-edges, as the reference does; they go after the invalidated vertices, counted apart
    // (cand_count stays fixed here). The engine's own tree needs less, see phase D.
    for (int64_t base = tid - lane(); p.check_cycles && base < p.num_ins; base += threads) {
      const int64_t i = base + lane();
      int32_t v = -1;
      bool add = false, first = false;
      if (i < p.num_ins) {
        v = p.ins_dst[i];
        add = v >= 0 && v != p.source && claim_first(p, v, generation, first);
      }
      append(p.cand + inv, &c->heads_count, add, v);
      append(p.touched, &c->touched_count, first, v);
    }
    DYNG_GRID_SYNC();
    // ---- D: pull ----
    ++generation;
    const int cands = inv + uniform(ld(&c->heads_count));
    if (tid == 0) {
      c->t_phase[1] = __builtin_amdgcn_s_memrealtime();
      c->jump_items = jump_items;
      c->pull_vertices = cands + (p.check_cycles ? 0 : p.num_ins);
    }
    int* chunk_count = &c->ring_chunk[cstep % 3];
    if (tid == 0) st(&c->ring_chunk[(cstep + 1) % 3], 0);
    if (!p.check_cycles) {
      // The engine's own tree is canonical: for a head v that was not invalidated, every in-edge
      // that was there before the batch and not reweighted offers at least v's old word (no
      // smaller distance, and no equal distance from a lower id), and v's old parent edge is
      // intact, so v's word can only fall through the edges this batch inserted or reweighted.
      // Each effective insertion (u, v) offers over the alive (u, v) entries of u's row, instead
      // of a pull over all of v's in-edges.
      for (int64_t base = tid - lane(); base < p.num_ins; base += threads) {
        const int64_t i = base + lane();
        int32_t v = -1;
        u64 best = packed_inf;
        if (i < p.num_ins) {
          v = p.ins_dst[i];
          if (v >= 0 && v != p.source) {
            best = p.ins_fw != nullptr ? final_offer(p, p.ins_src[i], i) : edge_offer(p, p.ins_src[i], v);
          }
        }
        pull_commit(p, best != packed_inf, v, best, generation, S, &c->frontier_count, &c->min_frontier);
      }
      stage_flush(S, p.frontier, &c->frontier_count);
    }
    tile_rows(
        L, cands,
        [&](int idx, int32_t& v, int32_t& off, int32_t& len, u64& word) {
          v = ld(&p.cand[idx]);
          const int2 mt = p.in_meta[v];
          word = packed_inf;  // the best offer so far
          if (mt.y <= p.light) {
            off = mt.x;
            len = mt.y;
            atomicAdd(pull_edges, static_cast<unsigned long long>(mt.y));
          } else {
            emit_chunks(p, chunk_count, v, mt.x, mt.y);
          }
        },
        [&](const bool (&act)[kSlots], const int (&item)[kSlots], const int (&slot)[kSlots]) {
          int32_t u[kSlots], wt[kSlots];
          u64 word[kSlots];
#pragma unroll
          for (int j = 0; j < kSlots; ++j) {
            u[j] = act[j] ? p.in_nbr[slot[j]] : -1;
            wt[j] = act[j] ? p.in_w[slot[j]] : 0;
          }
#pragma unroll
          for (int j = 0; j < kSlots; ++j) word[j] = u[j] >= 0 ? ld(&p.packed[u[j]]) : packed_inf;
#pragma unroll
          for (int j = 0; j < kSlots; ++j) {
            const u64 o = offer_of(p, u[j], wt[j], word[j]);
            if (act[j] && o < L.word[item[j]]) atomicMin(&L.word[item[j]], o);
          }
        },
        [&](bool valid, int item) {
          pull_commit(p, valid, valid ? L.vert[item] : -1, valid ? L.word[item] : packed_inf, generation, S,
                      &c->frontier_count, &c->min_frontier);
        },
        [&] { stage_flush(S, p.frontier, &c->frontier_count); });
    DYNG_GRID_SYNC();
    const int nchunks = uniform(static_cast<int>(min(static_cast<int64_t>(ld(chunk_count)), p.chunk_cap)));
    ++cstep;
    if (nchunks > 0) {
      const int wave = tid / 64;
      const int waves = threads / 64;
      for (int q = wave; q < nchunks; q += waves) {
        const int4 ch = ld_chunk(&p.chunks[q]);
        u64 best = packed_inf;
        for (int e = ch.y + lane(); e < ch.z; e += 64) {
          const u64 o = pull_offer(p, p.in_nbr[e], p.in_w[e]);
          best = o < best ? o : best;
        }
        if (lane() == 0) atomicAdd(pull_edges, static_cast<unsigned long long>(ch.z - ch.y));
        best = wave_min(best);
        if (lane() == 0) {
          const u64 current = ld(&p.packed[ch.x]);
          if (best < current) {
            const u64 old = atomicMin(&p.packed[ch.x], best);
            bool first = false;
            if (best < old && p.pk.distance(best) == p.pk.distance(old)) first = mark_changed(p, ch.x);
            if (p.pk.distance(best) < p.pk.distance(old)) {
              atomicMin(&c->min_frontier, p.pk.distance(best));
              if (claim_first(p, ch.x, generation, first)) {
                const int slot = atomicAdd(&c->frontier_count, 1);
                st(&p.frontier[slot], ch.x);
              }
            }
            if (first) st(&p.touched[atomicAdd(&c->touched_count, 1)], ch.x);
          }
        }
      }
      DYNG_GRID_SYNC();
    }
    frontier_count = uniform(ld(&c->frontier_count));
  }

  // ---- E: near-far propagation (the frontier's smallest distance came with the pull, or is the
  // source's 0) ----
  if (tid == 0) c->t_phase[2] = __builtin_amdgcn_s_memrealtime();
  u64 local = packed_inf;
  const u64 smallest = p.from_scratch ? 0 : uniform(ld(&c->min_frontier));
  u64 threshold = (smallest == packed_inf ? 0 : smallest) + p.delta;
  int* current = p.near_a;
  int* next = p.near_b;
  // the far piles: far_of(fb) is the current one (a select, not a two-entry array: a dynamically
  // indexed private array lives in scratch)
  auto far_of = [&](int b) { return b ? p.far_b : p.far_a; };
  int fb = 0;
  // The pull's frontier is the first relax iteration's list (each vertex once, claimed and
  // stamped by the pull): a vertex at or past the threshold goes to the far pile when its row
  // would be loaded, so no separate split pass and grid barrier come before the first iteration.
  const int* seed = p.frontier;
  int near_count = frontier_count;
  int iterations = 0, epochs = 0;
  long long pushes = 0;
  int pending = 0;  // chunks of long rows the previous relax iteration left for the next one
  // every relax iteration expands a vertex whose distance fell since its last expansion; a run
  // past this bound is a bug: stop with an internal error instead of spinning
  const long long guard = 4LL * n + 100000;
  for (;;) {
    if (static_cast<long long>(iterations) + epochs > guard) {
      if (tid == 0) atomicOr(&c->error_bits, DYNG_HIP_DEVICE_ERROR_INTERNAL);
      break;
    }
    if (near_count > 0 || pending > 0) {
      // -- one relax iteration, one grid barrier: rows up to `light` edges edge-balanced per
      // workgroup; a longer row is cut into chunks of chunk_edges, which every wave of the grid
      // relaxes at the start of the next iteration, before its tile (the chunks of iteration k
      // and k + 1 sit in the two halves of the chunk list; a vertex a chunk improves is claimed
      // with the next iteration's generation and expanded one iteration later, and the fixed
      // point, so the tree, is the same). After the first kRelaxGlobal iterations of a repair, an
      // iteration over a narrow near frontier lets each workgroup relax its share for up to
      // kRelaxLocal rounds on its own: the vertices it claims go to its next local level (LDS)
      // rather than to the next near list, each round claiming with a generation of its own
      // (generation0 + round * workgroups + workgroup), so a vertex improved after a workgroup
      // expanded it is claimed again and expanded again (every improvement is expanded, as in a
      // global iteration; the fixed point, and so the tree, is the same). A level past LDS and
      // the last level go to the next near list, once per vertex (the `ancestor` array, free in
      // this phase, tags them with the iteration). The deep, narrow repairs of a high-diameter
      // graph take hundreds of iterations; each global one costs a grid barrier and a list round
      // trip --
      // (a local iteration takes workgroups * kRelaxLocal generations: past kLocalGenerations
      // the iterations stay global, one generation each, so a run of millions of levels cannot
      // carry the stamps past INT32_MAX). An iteration with chunks to relax is global.
      const bool local = !p.from_scratch && iterations >= kRelaxGlobal && pending == 0 &&
                         near_count <= static_cast<int>(gridDim.x) * kRelaxNarrow && generation < kLocalGenerations;
      ++iterations;
      pushes += near_count;
      const int gen0 = generation + 1;
      // a global iteration claims with one generation; a local one reserves one per workgroup
      // and round
      generation += local ? static_cast<int>(gridDim.x) * kRelaxLocal : 1;
      int* next_count = &c->ring_near[step % 3];
      int* chunk_count = &c->ring_chunk[cstep % 3];
      const int64_t chunk_half = p.chunk_cap / 2;
      int4* chunk_out = p.chunks + (cstep & 1) * chunk_half;               // this iteration's chunks
      const int4* chunk_in = p.chunks + ((cstep + 1) & 1) * chunk_half;    // the previous one's
      if (tid == 0) {
        st(&c->ring_near[(step + 1) % 3], 0);
        st(&c->ring_chunk[(cstep + 1) % 3], 0);
      }
      int* far_list = far_of(fb);
      int* far_counter = &c->far_count[fb];
      const bool seeding = seed != nullptr;
      const int* cur = seeding ? seed : current;
      seed = nullptr;
      const int tag = local ? -2 - generation : 0;  // a local iteration's list tag: below every vertex id and -1
      auto load_row = [&](int32_t v, int32_t& off, int32_t& len, u64& word) {
        const int2 mt = p.out_meta[v];  // issued together with the word's load
        word = ld(&p.packed[v]);
        if (word == packed_inf) return;
        if (seeding && p.pk.distance(word) >= threshold) {
          // a vertex of the pull's frontier past the threshold waits in the far pile
          stage_append(S, 1, atomicExch(&p.in_far[v], 1) == 0, v, far_list, far_counter);
          return;
        }
        if (mt.y <= p.light) {
          off = mt.x;
          len = mt.y;
          atomicAdd(relax_edges, static_cast<unsigned long long>(mt.y));
        } else {
          emit_chunks_to(p, chunk_out, chunk_half, chunk_count, v, mt.x, mt.y);
        }
      };
      if (pending > 0) {
        // the previous iteration's chunks: one wave per chunk, grid-wide, then a flush (a tile
        // may not follow)
        const near_sink sink{nullptr, 0, nullptr, 0};
        const int wave = tid / 64;
        const int waves = threads / 64;
        for (int q = wave; q < pending; q += waves) {
          const int4 ch = ld_chunk(&chunk_in[q]);
          const int32_t u = ch.x;
          const u64 word = ld(&p.packed[u]);
          const u64 du = p.pk.distance(word);
          if (lane() == 0) atomicAdd(relax_edges, static_cast<unsigned long long>(ch.z - ch.y));
          for (int b = ch.y; b < ch.z; b += 64 * kSlots) {
            bool act[kSlots];
            int32_t uu[kSlots];
            u64 dd[kSlots];
            int slot[kSlots];
#pragma unroll
            for (int j = 0; j < kSlots; ++j) {
              const int e = b + j * 64 + lane();
              act[j] = word != packed_inf && e < ch.z;
              uu[j] = u;
              dd[j] = du;
              slot[j] = act[j] ? e : 0;
            }
            relax_slots(p, act, uu, dd, slot, threshold, generation, S, next, next_count, far_list, far_counter,
                        relax_improved, sink);
          }
        }
        stage_flush(S, next, next_count, far_list, far_counter);
      }
      if (!local) {
        const near_sink sink{nullptr, 0, nullptr, 0};
        tile_rows(
            L, near_count,
            [&](int idx, int32_t& v, int32_t& off, int32_t& len, u64& word) {
              v = ld(&cur[idx]);
              load_row(v, off, len, word);
            },
            [&](const bool (&act)[kSlots], const int (&item)[kSlots], const int (&slot)[kSlots]) {
              int32_t u[kSlots];
              u64 du[kSlots];
#pragma unroll
              for (int j = 0; j < kSlots; ++j) {
                u[j] = act[j] ? L.vert[item[j]] : 0;
                du[j] = act[j] ? p.pk.distance(L.word[item[j]]) : 0;
              }
              relax_slots(p, act, u, du, slot, threshold, generation, S, next, next_count, far_list, far_counter,
                          relax_improved, sink);
            },
            [&](bool, int) {}, [&] { stage_flush(S, next, next_count, far_list, far_counter); });
      } else {
        const int G = static_cast<int>(gridDim.x);
        const int ts = min(kBlock, max(1, (near_count + G - 1) / G));
        for (long long t0 = static_cast<long long>(blockIdx.x) * ts; t0 < near_count;
             t0 += static_cast<long long>(G) * ts) {
          const int cnt = static_cast<int>(min(static_cast<long long>(ts), near_count - t0));
          if (static_cast<int>(threadIdx.x) < cnt) W.item[0][threadIdx.x] = ld(&cur[t0 + threadIdx.x]);
          if (threadIdx.x == 0) {
            W.count[0] = cnt;
            W.count[1] = 0;
          }
          __syncthreads();
          int q = 0;
          int spent = 0;  // edges this tile's local rounds expanded
          for (int round = 0; round < kRelaxLocal && spent < kLocalEdgeBudget; ++round) {
            const int items = W.count[q];  // uniform: LDS after a workgroup barrier
            if (items == 0) break;
            if (round > 0 && threadIdx.x == 0) *local_pushes += static_cast<unsigned long long>(items);
            const int gen = gen0 + round * G + static_cast<int>(blockIdx.x);
            const near_sink sink{&W, q ^ 1, p.ancestor, tag};
            expand_tile(
                L, items,
                [&](int i, int32_t& v, int32_t& off, int32_t& len, u64& word) {
                  v = W.item[q][i];
                  load_row(v, off, len, word);
                },
                [&](const bool (&act)[kSlots], const int (&item)[kSlots], const int (&slot)[kSlots]) {
                  int32_t u[kSlots];
                  u64 du[kSlots];
#pragma unroll
                  for (int j = 0; j < kSlots; ++j) {
                    u[j] = act[j] ? L.vert[item[j]] : 0;
                    du[j] = act[j] ? p.pk.distance(L.word[item[j]]) : 0;
                  }
                  relax_slots(p, act, u, du, slot, threshold, gen, S, next, next_count, far_list, far_counter,
                              relax_improved, sink);
                },
                [&](bool, int) {});
            spent += L.start[items];  // the round's edges (LDS after the expansion's barrier: uniform)
            if (threadIdx.x == 0) {
              W.count[q] = 0;
              W.count[q ^ 1] = min(W.count[q ^ 1], kBlock);
            }
            if (S.count[0] > kStage / 2 || S.count[1] > kStage / 2 || S.count[2] > kStage / 2) {
              stage_flush(S, next, next_count, far_list, far_counter);  // uniform: LDS after a barrier
            }
            __syncthreads();
            q ^= 1;
          }
          // the level the last round left: the next near list, once per vertex
          const int left = W.count[q];
          for (int base = static_cast<int>(threadIdx.x) - lane(); base < left; base += kBlock) {
            const int i = base + lane();
            const int32_t v = i < left ? W.item[q][i] : -1;
            const bool add = i < left && atomicExch(&p.ancestor[v], tag) != tag;
            stage_append(S, 0, add, v, next, next_count);
          }
          stage_flush(S, next, next_count, far_list, far_counter);
        }
      }
      DYNG_GRID_SYNC();
      pending = uniform(static_cast<int>(min(static_cast<int64_t>(ld(chunk_count)), chunk_half)));
      ++cstep;
      near_count = uniform(ld(next_count));
      ++step;
      int* t = current;
      current = next;
      next = t;
      continue;
    }
    // -- near frontier empty: raise the threshold past the far pile --
    const int far_count = uniform(ld(&c->far_count[fb]));
    if (far_count == 0) break;
    ++epochs;
    const int ms = epochs & 1;
    if (tid == 0) st(&c->minimum[ms ^ 1], packed_inf);
    local = packed_inf;
    for (int i = tid; i < far_count; i += threads) {
      const u64 word = ld(&p.packed[ld(&far_of(fb)[i])]);
      if (word != packed_inf) local = min(local, p.pk.distance(word));
    }
    local = wave_min(local);
    if (lane() == 0 && local != packed_inf) atomicMin(&c->minimum[ms], local);
    DYNG_GRID_SYNC();
    const u64 mn = uniform(ld(&c->minimum[ms]));
    threshold = (threshold > mn ? threshold : mn) + p.delta;
    ++generation;
    int* split_count = &c->ring_near[step % 3];
    if (tid == 0) st(&c->ring_near[(step + 1) % 3], 0);
    for (int base = tid - lane(); base < far_count; base += threads) {
      const int i = base + lane();
      int32_t v = -1;
      bool to_near = false, keep = false, first = false;
      if (i < far_count) {
        v = ld(&far_of(fb)[i]);
        const u64 word = ld(&p.packed[v]);
        const u64 d = word == packed_inf ? packed_inf : p.pk.distance(word);
        const bool near = d < threshold;
        if (near) st(&p.in_far[v], 0);
        to_near = near && claim_first(p, v, generation, first);
        keep = !near;
      }
      stage_append(S, 0, to_near, v, current, split_count);
      stage_append(S, 1, keep, v, far_of(fb ^ 1), &c->far_count[fb ^ 1]);
      stage_append(S, 2, first, v, S.list2, S.counter2);
    }
    stage_flush(S, current, split_count, far_of(fb ^ 1), &c->far_count[fb ^ 1]);
    DYNG_GRID_SYNC();
    near_count = uniform(ld(split_count));
    ++step;
    if (tid == 0) st(&c->far_count[fb], 0);
    fb ^= 1;
  }

  // ---- F: unpack ----
  if (tid == 0) c->t_phase[3] = __builtin_amdgcn_s_memrealtime();
  unsigned long long changed = 0;
  if (p.pk.bits > 0) {
    // A repair changed only words of vertices it stamped (see mark_changed); every other word
    // still equals its vertex's packed (distance, parent), so only the stamped ones are visited,
    // from the lists that stamped them (no pass over every vertex). A run from the sources visits
    // every vertex.
    auto unpack = [&](int32_t v) {
      const u64 word = ld(&p.packed[v]);
      const long long d = word == packed_inf ? sssp_infinity : static_cast<long long>(p.pk.distance(word));
      const int32_t par = word == packed_inf ? -1 : p.pk.parent(word);
      if (!p.count_changes) {
        p.distances[v] = d;
        p.parents[v] = par;
      } else if (p.distances[v] != d || p.parents[v] != par) {
        p.distances[v] = d;
        p.parents[v] = par;
        ++changed;
      }
    };
    if (p.from_scratch) {
      for (int v = tid; v < n; v += threads) unpack(v);
    } else {
      // the invalidated vertices (cand[0, invalidated)) and the vertices whose first stamp of
      // this run came from a claim or a same-distance lowering (the touched list): disjoint, as
      // the invalidated ones were stamped before any claim
      const int inv = uniform(ld(&c->invalidated));
      const int total = inv + uniform(ld(&c->touched_count));
      for (int i = tid; i < total; i += threads) unpack(i < inv ? ld(&p.cand[i]) : ld(&p.touched[i - inv]));
    }
  } else {
    // distance-only words: lowest-id tight parents (fused.cuh:452-485)
    for (int v = tid; v < n; v += threads) {
      const u64 word = ld(&p.packed[v]);
      const long long d = word == packed_inf ? sssp_infinity : static_cast<long long>(word);
      if (p.count_changes) {
        p.ancestor[v] = p.parents[v];
        p.flag[v] = d != p.distances[v] ? 1 : 0;
      }
      p.distances[v] = d;
      p.parents[v] = (word == packed_inf || word == 0) ? -1 : INT32_MAX;  // distance 0: a source
    }
    DYNG_GRID_SYNC_FENCED();
    for (int u = tid; u < n; u += threads) {
      const u64 du = ld(&p.packed[u]);
      if (du == packed_inf) continue;
      const int2 mt = p.out_meta[u];
      for (int j = 0; j < mt.y; ++j) {
        const int32_t w = p.out_nbr[mt.x + j];
        if (w >= 0 && w != p.source && du + static_cast<u64>(p.out_w[mt.x + j]) == ld(&p.packed[w])) atomicMin(&p.parents[w], u);
      }
    }
    if (p.count_changes) {
      DYNG_GRID_SYNC_FENCED();
      for (int v = tid; v < n; v += threads) {
        changed += (p.flag[v] != 0 || p.ancestor[v] != ld(&p.parents[v])) ? 1 : 0;
        p.flag[v] = 0;
      }
    }
  }
  changed = wave_sum(changed);
  if (lane() == 0 && changed) atomicAdd(affected, changed);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (*affected) atomicAdd(&c->affected, *affected);
    if (*relax_edges) atomicAdd(&c->relax_edges, *relax_edges);
    if (*relax_improved) atomicAdd(&c->relax_improved, *relax_improved);
    if (*pull_edges) atomicAdd(&c->pull_edges, *pull_edges);
    if (*local_pushes) atomicAdd(reinterpret_cast<unsigned long long*>(&c->pushes), *local_pushes);
  }
  if (tid == 0) {
    c->t_phase[7] = __builtin_amdgcn_s_memrealtime();
    c->iterations = iterations;
    c->epochs = epochs;
    atomicAdd(reinterpret_cast<unsigned long long*>(&c->pushes), static_cast<unsigned long long>(pushes));
    c->generation = generation;
  }
}

#undef DYNG_GRID_SYNC
#undef DYNG_GRID_SYNC_FENCED
#undef DYNG_GRID_SYNC_AS

}  // namespace dyng_hip

// ---------------------------------------------------------------------------------------------
// Workspace and launch
// ---------------------------------------------------------------------------------------------

struct dyng_hip_workspace {
  int device = 0;
  int64_t capacity = 0;
  int generation = 0;
  dyng_hip::dbuf<dyng_hip::u64> packed;
  dyng_hip::dbuf<int> stamp, in_far, flag, ancestor;
  dyng_hip::dbuf<int> cand, frontier, near_a, near_b, far_a, far_b, touched;
  /// the tree whose packed words `packed` holds after the last successful run: its token (0:
  /// none), size, parent bits and distance array
  uint64_t words_token = 0;
  int64_t words_n = 0;
  int words_bits = 0;
  const void* words_distances = nullptr;
  dyng_hip::dbuf<int4> chunks;
  int64_t chunk_cap = 0;
  dyng_hip::dbuf<dyng_hip::control> ctl;
  dyng_hip::pinned<dyng_hip::control> host_ctl;
  dyng_hip::dbuf<dyng_hip::grid_sync_state> bar;
  int grid_blocks = 0;
  int cu_count = 0;
  int occupancy = 0;         ///< workgroups per CU of the 64-VGPR build
  int occupancy_narrow = 0;  ///< and of the 128-VGPR build
  hipEvent_t ev_start = nullptr;
  hipEvent_t ev_stop = nullptr;
  /// the launch as a hipGraph (control upload, barrier reset, events, the kernel, control
  /// download), built once and replayed with the kernel node's arguments updated
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  hipGraphNode_t graph_kernel = nullptr;
  bool graph_narrow = false;  ///< which build the graph's kernel node launches
  ~dyng_hip_workspace() {
    if (graph_exec != nullptr) (void)hipGraphExecDestroy(graph_exec);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    if (ev_start != nullptr) (void)hipEventDestroy(ev_start);
    if (ev_stop != nullptr) (void)hipEventDestroy(ev_stop);
  }
};

namespace dyng_hip {

namespace {

/// The build of the persistent kernel a grid takes: the 128-VGPR one when the grid fits at its
/// occupancy (one workgroup per CU), else the 64-VGPR one.
const void* fused_kernel(bool narrow) {
  return narrow ? reinterpret_cast<const void*>(sssp_fused_kernel<4>) : reinterpret_cast<const void*>(sssp_fused_kernel<8>);
}
bool narrow_grid(int blocks, int occupancy_narrow, int cu_count) {
  return occupancy_narrow >= 1 && blocks <= occupancy_narrow * cu_count;
}

/// Chunk-list entries a run over n vertices and a row arena of `slots` slots may need: one pass
/// over every row in chunks of at least 16 edges takes at most slots / 16 + n; the relax
/// iterations keep two such lists (one iteration's chunks are relaxed during the next).
int64_t chunks_for(int64_t n, int64_t slots) { return 2 * (slots / 16 + n + 4096); }

void ws_reserve(dyng_hip_workspace* ws, int64_t n, int64_t slots, hipStream_t s) {
  const int64_t want_chunks = chunks_for(n, slots);
  if (n > ws->capacity) {
    const int64_t cap = std::max<int64_t>(n, 1);
    ws->packed.reserve(cap);
    ws->stamp.reserve(cap);
    ws->in_far.reserve(cap);
    ws->flag.reserve(cap);
    ws->ancestor.reserve(cap);
    ws->cand.reserve(2 * cap + 64);
    ws->frontier.reserve(cap);
    ws->near_a.reserve(cap);
    ws->near_b.reserve(cap);
    ws->far_a.reserve(cap);
    ws->far_b.reserve(cap);
    ws->touched.reserve(cap);
    DYNG_HIP_TRY(hipMemsetAsync(ws->stamp.p, 0, cap * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->in_far.p, 0, cap * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->flag.p, 0, cap * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->ancestor.p, 0, cap * sizeof(int), s));
    ws->capacity = cap;
    ws->generation = 0;
    ws->words_token = 0;
  }
  if (want_chunks > ws->chunk_cap) {
    ws->chunks.reserve(want_chunks);
    ws->chunk_cap = want_chunks;
  }
  ws->ctl.reserve(1);
  ws->bar.reserve(1);
  if (ws->cu_count == 0) {
    DYNG_HIP_TRY(hipDeviceGetAttribute(&ws->cu_count, hipDeviceAttributeMultiprocessorCount, ws->device));
    DYNG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&ws->occupancy, fused_kernel(false), kBlock, 0));
    DYNG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&ws->occupancy_narrow, fused_kernel(true), kBlock, 0));
    if (ws->occupancy < 1) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: sssp: the persistent kernel does not fit on a CU");
  }
}

int bits_for_parents(int64_t n) {
  int bits = 1;
  while (((1ULL << bits) - 1) < static_cast<u64>(n)) ++bits;
  return bits;
}

}  // namespace

namespace {

/// Where a run's scratch lives: the device arrays, the control block and the barrier state.
struct engine_scratch {
  u64* packed;
  int *stamp, *in_far, *flag, *ancestor;
  int *cand, *frontier, *near_a, *near_b, *far_a, *far_b, *touched;
  int4* chunks;
  int64_t chunk_cap;
  control* ctl;
  grid_sync_state* bar;
};

/// The checks and the parameters of a run that do not depend on where its scratch lives (a
/// workspace, or the no-allocation entry point's static array). n > 0.
void engine_setup(const dyng_hip_sssp_run* run, bool update, engine_params& p, int64_t& timeout_us) {
  const dyng_hip_graph* g = run->graph;
  const int64_t n = g->n;
  if (run->delta <= 0) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: the near-far width delta must be > 0, got " + std::to_string(run->delta));
  if (run->source == -1) {
    if (run->sources == nullptr || run->num_sources < 1 || run->num_sources > n) {
      fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: source -1 needs a device list of 1 to n sources");
    }
  } else if (run->source < 0 || run->source >= n) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: source " + std::to_string(run->source) + " is out of range [0, " +
                                        std::to_string(n) + ")");
  }
  if (run->objective < 0 || run->objective >= g->K) fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: objective out of range");
  if (run->tuning.engine == DYNG_HIP_ENGINE_OPERATORS) {
    fail(DYNG_HIP_NOT_SUPPORTED, "dyng: sssp: the operators engine is not built in this release; use fused");
  }
  // choose_packing (cuda.cu:226-234): (n - 1) * max_weight must fit next to the parent bits
  const u64 weight = static_cast<u64>(std::max<int64_t>(run->max_weight, 1));
  const u64 hops = static_cast<u64>(std::max<int64_t>(n - 1, 1));
  if (weight > output_max_distance / hops) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "sssp: distances up to " + std::to_string(weight) + " * " + std::to_string(hops) +
                                        " do not fit in 62 bits");
  }
  const u64 bound = weight * hops;
  const int bits = bits_for_parents(n);
  const u64 max_packed = (packed_inf >> bits) - 1;
  packing pk{};
  if (bound <= max_packed) {
    pk.bits = bits;
    pk.no_parent = (1ULL << bits) - 1;
  } else {
    pk.bits = 0;
    pk.no_parent = 0;
  }
  p = engine_params{};
  p.n = static_cast<int32_t>(n);
  const int k = run->objective;
  p.out_meta = g->out.meta.p;
  p.out_base = g->out.base.p;
  p.out_nbr = g->out.nbr.p;
  p.out_w = g->out.w.p + static_cast<int64_t>(k) * g->out.slot_cap;
  p.in_meta = g->in.meta.p;
  p.in_nbr = g->in.nbr.p;
  p.in_w = g->in.w.p + static_cast<int64_t>(k) * g->in.slot_cap;
  if (update) {
    p.del_src = g->del_src.p;
    p.del_dst = g->del_dst.p;
    p.num_del = g->num_del_slots;
    p.ins_src = g->ins_src.p;
    p.ins_dst = g->ins_dst.p;
    p.ins_inc = g->ins_inc.p;
    p.ins_fw = g->props.allow_parallel ? nullptr : g->ins_fw.p;
    p.inc_stride = g->K;
    p.inc_k = k;
    p.num_ins = g->num_ins_slots;
  }
  p.source = run->source;
  p.sources = run->sources;
  p.num_sources = run->num_sources;
  // an update may recompute from the sources instead of repairing (the same tree; `affected`
  // still counts the changed pairs against the old tree)
  const bool recompute = update && run->tuning.recompute != 0;
  p.from_scratch = update && !recompute ? 0 : 1;
  p.count_changes = update ? 1 : 0;
  p.check_cycles = run->check_cycles;
  p.pk = pk;
  p.max_distance = bound;
  p.delta = static_cast<u64>(run->delta);
  int rounds = 0;
  while ((1LL << rounds) < static_cast<long long>(n)) ++rounds;
  p.max_rounds = rounds + 1;
  p.distances = reinterpret_cast<long long*>(run->distances);
  p.parents = run->parents;
  // rows up to 8192 edges in a repair's tiles (4096 measured 6 % slower on RMAT-22 repairs: more
  // rows waited a relax iteration as chunks), up to 4096 in a solve's (its wide frontiers balance
  // better with the long rows in chunks: 8192 measured 5.9 -> 8.7 ms on RMAT-22)
  p.light = run->tuning.light_degree > 0 ? run->tuning.light_degree : (p.from_scratch ? 4096 : 8192);
  p.chunk_edges = run->tuning.chunk_edges > 0 ? std::max(16, run->tuning.chunk_edges) : 256;
  // the real-time counter runs at 100 MHz: 100 ticks per microsecond
  timeout_us = run->tuning.barrier_timeout_us > 0   ? run->tuning.barrier_timeout_us
               : run->tuning.barrier_timeout_us == 0 ? 4000000
                                                     : 0;
  p.barrier_ticks = static_cast<unsigned long long>(timeout_us) * 100ULL;
}

void attach_scratch(engine_params& p, const engine_scratch& sc) {
  p.packed = sc.packed;
  p.stamp = sc.stamp;
  p.in_far = sc.in_far;
  p.flag = sc.flag;
  p.ancestor = sc.ancestor;
  p.cand = sc.cand;
  p.frontier = sc.frontier;
  p.near_a = sc.near_a;
  p.near_b = sc.near_b;
  p.far_a = sc.far_a;
  p.far_b = sc.far_b;
  p.touched = sc.touched;
  p.chunks = sc.chunks;
  p.chunk_cap = sc.chunk_cap;
  p.ctl = sc.ctl;
  p.bar = sc.bar;
}

/// Every workgroup must be resident for the grid barrier. A solve from the sources alone on the
/// device fills it (two 1024-thread workgroups per CU: on its wide frontiers the waves of one hide
/// the other's workgroup barriers and gathers, 1.6x on RMAT-22 against one per CU); a repair, whose
/// frontiers are narrow and whose steps are chains of dependent accesses, runs one per CU (measured
/// 4 % faster on RMAT-22 batches of 100,000 operations than two). Runs never overlap on a device
/// (they are queued on one stream). The barrier's bounded wait turns a grid that is not resident
/// after all into an error.
int grid_blocks(const dyng_hip_tuning& t, int occupancy, int cu_count, bool from_sources) {
  const int resident = std::max(1, occupancy * cu_count);  // what the runtime says fits at once
  // an explicit count may go up to everything that fits (the caller runs nothing else beside it)
  if (t.grid_blocks > 0) return std::min(t.grid_blocks, resident);
  const int per_cu = t.blocks_per_cu > 0 ? t.blocks_per_cu : (!from_sources ? 1 : occupancy);
  return std::max(1, std::min(per_cu * cu_count, resident));
}

/// The errors a run's control block reports, as the reference's device errors.
void throw_on_errors(const control* hc, int64_t timeout_us) {
  if (hc->overflow != 0) {
    throw status_error(DYNG_HIP_DEVICE_ERROR,
                       "sssp: an input distance is negative or larger than (n - 1) * max weight; the initial tree "
                       "does not belong to this graph",
                       DYNG_HIP_DEVICE_ERROR_INVALID_INPUT);
  }
  if (hc->error_bits & DYNG_HIP_DEVICE_ERROR_TIMEOUT) {
    throw status_error(DYNG_HIP_DEVICE_ERROR,
                       "sssp: a grid barrier wait of the persistent kernel ran past " + std::to_string(timeout_us) +
                           " us (its workgroups were not all resident: is other work filling the device?)",
                       hc->error_bits);
  }
  if (hc->error_bits & DYNG_HIP_DEVICE_ERROR_PARENT_CYCLE) {
    throw status_error(DYNG_HIP_DEVICE_ERROR,
                       "sssp: the input shortest-path tree has a parent cycle (pointer jumping did not reach a root "
                       "in ceil(log2 n) + 1 rounds)",
                       DYNG_HIP_DEVICE_ERROR_PARENT_CYCLE);
  }
  if (hc->error_bits != 0) {
    throw status_error(DYNG_HIP_DEVICE_ERROR, "sssp: the device reported error bits " + std::to_string(hc->error_bits),
                       hc->error_bits);
  }
}

/// The counters of a finished run from its control block.
void read_counters(const control* hc, bool update, int bits, int blocks, dyng_hip_counters& counters) {
  counters.invalidated = hc->invalidated;
  counters.iterations = hc->iterations;
  counters.epochs = hc->epochs;
  counters.pushes = hc->pushes;
  counters.affected = static_cast<int64_t>(hc->affected);
  counters.packed_parents = bits > 0 ? 1 : 0;
  counters.relax_vertices = hc->pushes;
  counters.relax_edges = static_cast<int64_t>(hc->relax_edges);
  counters.relax_improved = static_cast<int64_t>(hc->relax_improved);
  counters.pull_edges = static_cast<int64_t>(hc->pull_edges);
  counters.jump_rounds = hc->rounds;
  counters.jump_items = static_cast<int64_t>(hc->jump_items);
  counters.pull_vertices = static_cast<int64_t>(hc->pull_vertices);
  // s_memrealtime runs at 100 MHz
  const double tick_ms = 1e-5;
  const unsigned long long t0 = hc->t_phase[0];
  const unsigned long long t1 = update ? hc->t_phase[1] : t0;
  const unsigned long long t2 = hc->t_phase[2];
  const unsigned long long t3 = hc->t_phase[3];
  counters.step1_ms = (t2 >= t0 ? t2 - t0 : 0) * tick_ms;
  counters.pull_ms = (update && t2 >= t1 ? t2 - t1 : 0) * tick_ms;
  counters.relax_ms = (t3 >= t2 ? t3 - t2 : 0) * tick_ms;
  const unsigned long long ta = hc->t_phase[4], tb = hc->t_phase[5], t7 = hc->t_phase[7];
  counters.pack_ms = update && ta >= t0 ? (ta - t0) * tick_ms : 0.0;
  counters.jump_ms = update && tb >= ta ? (tb - ta) * tick_ms : 0.0;
  counters.unpack_ms = t7 >= t3 ? (t7 - t3) * tick_ms : 0.0;
  counters.blocks = blocks;
}

}  // namespace

engine_pending engine_launch(const dyng_hip_sssp_run* run, bool update, hipStream_t s, const int* hold) {
  if (run == nullptr || run->graph == nullptr || run->workspace == nullptr) {
    fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: sssp run: graph and workspace are required");
  }
  const dyng_hip_graph* g = run->graph;
  dyng_hip_workspace* ws = run->workspace;
  const int64_t n = g->n;
  engine_pending pend{};
  pend.ws = ws;
  pend.s = s;
  pend.update = update;
  pend.n = n;
  pend.distances = run->distances;
  dyng_hip_counters& counters = pend.counters;
  counters.engine_used = DYNG_HIP_ENGINE_FUSED;
  counters.packed_parents = 1;
  if (n == 0) {
    pend.launched = false;
    return pend;
  }
  engine_params p{};
  int64_t timeout_us = 0;
  engine_setup(run, update, p, timeout_us);
  p.hold = hold;
  ws_reserve(ws, n, std::max(g->out.slot_cap, g->in.slot_cap), s);
  if (ws->generation > INT32_MAX / 2) {
    // stamps and the relax phase's list tags (-2 - generation, in `ancestor`) restart
    DYNG_HIP_TRY(hipMemsetAsync(ws->stamp.p, 0, ws->capacity * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->ancestor.p, 0, ws->capacity * sizeof(int), s));
    ws->generation = 0;
  }
  p.generation = ++ws->generation;
  // the workspace still holds this tree's words when the caller's token is the one the last
  // successful run handed out for them (same size, packing and arrays)
  const bool recompute = update && run->tuning.recompute != 0;
  p.resident = update && !recompute && !run->check_cycles && p.pk.bits > 0 && run->tree_token != 0 &&
                       run->tree_token == ws->words_token && ws->words_n == n && ws->words_bits == p.pk.bits &&
                       ws->words_distances == run->distances
                   ? 1
                   : 0;
  ws->words_token = 0;  // until this run ends well
  attach_scratch(p, engine_scratch{ws->packed.p, ws->stamp.p, ws->in_far.p, ws->flag.p, ws->ancestor.p, ws->cand.p,
                                   ws->frontier.p, ws->near_a.p, ws->near_b.p, ws->far_a.p, ws->far_b.p,
                                   ws->touched.p, ws->chunks.p, ws->chunk_cap, ws->ctl.p, ws->bar.p});
  const int blocks = grid_blocks(run->tuning, ws->occupancy, ws->cu_count, p.from_scratch != 0);
  const bool narrow = narrow_grid(blocks, ws->occupancy_narrow, ws->cu_count);
  control* hc = ws->host_ctl.p;
  std::memset(hc, 0, sizeof(control));
  hc->min_frontier = packed_inf;
  hc->minimum[0] = packed_inf;
  hc->minimum[1] = packed_inf;
  if (ws->ev_start == nullptr) {
    DYNG_HIP_TRY(hipEventCreate(&ws->ev_start));
    DYNG_HIP_TRY(hipEventCreate(&ws->ev_stop));
  }
  void* args[] = {&p};
  if (run->tuning.launch == DYNG_HIP_LAUNCH_GRAPH) {
    hipKernelNodeParams kp{};
    kp.func = const_cast<void*>(fused_kernel(narrow));
    kp.gridDim = dim3(blocks);
    kp.blockDim = dim3(kBlock);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    if (ws->graph_exec != nullptr && ws->graph_narrow != narrow) {
      // the other build of the kernel: the graph is built again around it
      DYNG_HIP_TRY(hipGraphExecDestroy(ws->graph_exec));
      DYNG_HIP_TRY(hipGraphDestroy(ws->graph));
      ws->graph_exec = nullptr;
      ws->graph = nullptr;
    }
    if (ws->graph_exec == nullptr) {
      // built once per workspace: the control block, the barrier and the events never move
      ws->graph_narrow = narrow;
      DYNG_HIP_TRY(hipGraphCreate(&ws->graph, 0));
      hipGraphNode_t reset, start, stop, down;
      grid_sync_state* bar = ws->bar.p;
      control* ctl = ws->ctl.p;
      void* reset_args[] = {&bar, &ctl};
      hipKernelNodeParams rp{};
      rp.func = reinterpret_cast<void*>(k_reset_barrier);
      rp.gridDim = dim3(1);
      rp.blockDim = dim3(64);
      rp.kernelParams = reset_args;
      DYNG_HIP_TRY(hipGraphAddKernelNode(&reset, ws->graph, nullptr, 0, &rp));
      DYNG_HIP_TRY(hipGraphAddEventRecordNode(&start, ws->graph, &reset, 1, ws->ev_start));
      DYNG_HIP_TRY(hipGraphAddKernelNode(&ws->graph_kernel, ws->graph, &start, 1, &kp));
      DYNG_HIP_TRY(hipGraphAddEventRecordNode(&stop, ws->graph, &ws->graph_kernel, 1, ws->ev_stop));
      DYNG_HIP_TRY(hipGraphAddMemcpyNode1D(&down, ws->graph, &stop, 1, hc, ws->ctl.p, sizeof(control),
                                           hipMemcpyDeviceToHost));
      DYNG_HIP_TRY(hipGraphInstantiate(&ws->graph_exec, ws->graph, nullptr, nullptr, 0));
    } else {
      DYNG_HIP_TRY(hipGraphExecKernelNodeSetParams(ws->graph_exec, ws->graph_kernel, &kp));
    }
    DYNG_HIP_TRY(hipGraphLaunch(ws->graph_exec, s));
  } else {
    hipLaunchKernelGGL(k_reset_barrier, dim3(1), dim3(64), 0, s, ws->bar.p, ws->ctl.p);
    DYNG_HIP_TRY(hipEventRecord(ws->ev_start, s));
    if (narrow) {
      hipLaunchKernelGGL(sssp_fused_kernel<4>, dim3(blocks), dim3(kBlock), 0, s, p);
    } else {
      hipLaunchKernelGGL(sssp_fused_kernel<8>, dim3(blocks), dim3(kBlock), 0, s, p);
    }
    DYNG_HIP_TRY(hipGetLastError());
    DYNG_HIP_TRY(hipEventRecord(ws->ev_stop, s));
    DYNG_HIP_TRY(hipMemcpyAsync(hc, ws->ctl.p, sizeof(control), hipMemcpyDeviceToHost, s));
  }
  pend.launched = true;
  pend.bits = p.pk.bits;
  pend.blocks = blocks;
  pend.resident = p.resident;
  pend.timeout_us = timeout_us;
  return pend;
}

void engine_finish(engine_pending& pend, dyng_hip_counters* out) {
  dyng_hip_counters counters = pend.counters;
  if (!pend.launched) {
    if (out) *out = counters;
    return;
  }
  pend.launched = false;
  dyng_hip_workspace* ws = pend.ws;
  hipStream_t s = pend.s;
  const bool update = pend.update;
  const int64_t timeout_us = pend.timeout_us;
  control* hc = ws->host_ctl.p;
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (hc->held != 0) {
    // nothing ran (the apply ahead of the run was held back): the caller runs it again
    pend.held = true;
    if (out) *out = counters;
    return;
  }
  float kernel_ms = 0.0f;
  DYNG_HIP_TRY(hipEventElapsedTime(&kernel_ms, ws->ev_start, ws->ev_stop));
  if (hc->overflow != 0 || hc->error_bits != 0) {
    // a run that stopped early can leave flags, far marks and stamps of any later generation
    // behind in the shared workspace: clear them before any other run uses it
    DYNG_HIP_TRY(hipMemsetAsync(ws->stamp.p, 0, ws->capacity * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->in_far.p, 0, ws->capacity * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->flag.p, 0, ws->capacity * sizeof(int), s));
    DYNG_HIP_TRY(hipMemsetAsync(ws->ancestor.p, 0, ws->capacity * sizeof(int), s));
    DYNG_HIP_TRY(hipStreamSynchronize(s));
    ws->generation = 0;
  }
  throw_on_errors(hc, timeout_us);
  ws->generation = std::max(ws->generation, hc->generation);
  read_counters(hc, update, pend.bits, pend.blocks, counters);
  counters.kernel_ms = kernel_ms;
  counters.words_kept = pend.resident;
  if (pend.bits > 0) {
    // `packed` now holds the words of the tree the run left: hand out a token for them
    static std::atomic<uint64_t> next_token{0};
    ws->words_token = ++next_token;
    ws->words_n = pend.n;
    ws->words_bits = pend.bits;
    ws->words_distances = pend.distances;
    counters.tree_token = ws->words_token;
  }
  if (out) *out = counters;
}

namespace {

void run_engine(const dyng_hip_sssp_run* run, bool update, hipStream_t s, dyng_hip_counters* out) {
  engine_pending pend = engine_launch(run, update, s);
  engine_finish(pend, out);
}

}  // namespace

}  // namespace dyng_hip

// ---------------------------------------------------------------------------------------------
// The update that allocates nothing: its scratch is one static device array
// ---------------------------------------------------------------------------------------------

/// The scratch of dyng_hip_sssp_update_noalloc: the loader places one copy in each device's memory
/// when the code object loads there, and it stays until the process ends. Host code never passes
/// this name to a kernel or a copy (on the host it only stands for the symbol): it uses the device
/// address hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[DYNG_HIP_NOALLOC_SCRATCH_BYTES];

namespace dyng_hip {

namespace {

__global__ void k_clear_scratch(unsigned long long* p, int64_t words) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
}

void clear_scratch(void* p, int64_t bytes, hipStream_t s) {
  const int64_t words = bytes / 8;  // every region is a multiple of 256 bytes
  if (words <= 0) return;
  const int64_t groups = std::min<int64_t>((words + 255) / 256, 2048);
  hipLaunchKernelGGL(k_clear_scratch, dim3(static_cast<unsigned>(groups)), dim3(256), 0, s,
                     static_cast<unsigned long long*>(p), words);
  DYNG_HIP_TRY(hipGetLastError());
}

/// Where the arrays of a run over n vertices (row arenas of `slots` slots) sit in the scratch:
/// byte offsets, each region rounded up to 256 bytes.
struct scratch_layout {
  int64_t packed, stamp, in_far, flag, ancestor, cand, frontier, near_a, near_b, far_a, far_b, touched, chunks, ctl,
      bar;
  int64_t chunk_cap;
  int64_t total;
};

scratch_layout layout_for(int64_t n, int64_t slots) {
  scratch_layout L{};
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    const int64_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  L.packed = take(8 * n);
  L.stamp = take(4 * n);
  L.in_far = take(4 * n);
  L.flag = take(4 * n);
  L.ancestor = take(4 * n);
  L.cand = take(4 * (2 * n + 64));
  L.frontier = take(4 * n);
  L.near_a = take(4 * n);
  L.near_b = take(4 * n);
  L.far_a = take(4 * n);
  L.far_b = take(4 * n);
  L.touched = take(4 * n);
  L.chunk_cap = chunks_for(n, slots);
  L.chunks = take(static_cast<int64_t>(sizeof(int4)) * L.chunk_cap);
  L.ctl = take(sizeof(control));
  L.bar = take(sizeof(grid_sync_state));
  L.total = off;
  return L;
}

/// What the host keeps per device for the static scratch: its device address, the stamps'
/// generation, the launch shape, and the control block read back into plain host memory.
struct noalloc_device {
  bool ready = false;
  unsigned char* ptr = nullptr;
  int generation = 0;
  int64_t last_n = 0;  ///< the vertex count the stamps, far marks and flags were laid out for
  int cu_count = 0;
  int occupancy = 0;
  int occupancy_narrow = 0;
  control host_ctl{};
};

constexpr int kNoallocDevices = 64;
noalloc_device noalloc_state[kNoallocDevices];
std::mutex noalloc_setup_lock;

/// Once per device, with the device current: the scratch's address on it (hipGetSymbolAddress),
/// cleared once, so stamps, far marks and flags start at zero; and the launch shape.
noalloc_device& noalloc_for(int device, hipStream_t s) {
  if (device < 0 || device >= kNoallocDevices) {
    fail(DYNG_HIP_NOT_SUPPORTED, "dyng: sssp update (no allocation): device " + std::to_string(device) +
                                     " is past the " + std::to_string(kNoallocDevices) + " devices it serves");
  }
  noalloc_device& st = noalloc_state[device];
  std::lock_guard<std::mutex> hold(noalloc_setup_lock);
  if (st.ready) return st;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr) {
    (void)hipGetLastError();
    fail(DYNG_HIP_RUNTIME_ERROR, std::string("dyng: sssp update (no allocation): hipGetSymbolAddress of the scratch "
                                             "array failed: ") +
                                     hipGetErrorString(e));
  }
  DYNG_HIP_TRY(hipDeviceGetAttribute(&st.cu_count, hipDeviceAttributeMultiprocessorCount, device));
  DYNG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&st.occupancy, fused_kernel(false), kBlock, 0));
  DYNG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&st.occupancy_narrow, fused_kernel(true), kBlock, 0));
  if (st.occupancy < 1) fail(DYNG_HIP_NOT_SUPPORTED, "dyng: sssp: the persistent kernel does not fit on a CU");
  clear_scratch(ptr, DYNG_HIP_NOALLOC_SCRATCH_BYTES, s);
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  st.ptr = static_cast<unsigned char*>(ptr);
  st.generation = 0;
  st.ready = true;
  return st;
}

void run_engine_noalloc(const dyng_hip_sssp_run* run, hipStream_t s, dyng_hip_counters* out) {
  if (run == nullptr || run->graph == nullptr) fail(DYNG_HIP_INVALID_ARGUMENT, "dyng: sssp run: a graph is required");
  const dyng_hip_graph* g = run->graph;
  const int64_t n = g->n;
  dyng_hip_counters counters{};
  counters.engine_used = DYNG_HIP_ENGINE_FUSED;
  counters.packed_parents = 1;
  if (n == 0) {
    if (out) *out = counters;
    return;
  }
  engine_params p{};
  int64_t timeout_us = 0;
  engine_setup(run, true, p, timeout_us);
  const scratch_layout L = layout_for(n, std::max(g->out.slot_cap, g->in.slot_cap));
  if (L.total > static_cast<int64_t>(DYNG_HIP_NOALLOC_SCRATCH_BYTES)) {
    fail(DYNG_HIP_INVALID_ARGUMENT,
         "dyng: sssp update (no allocation): a graph of " + std::to_string(n) + " vertices and " +
             std::to_string(std::max(g->out.slot_cap, g->in.slot_cap)) + " row slots needs " +
             std::to_string(L.total) + " bytes of scratch, more than the " +
             std::to_string(static_cast<int64_t>(DYNG_HIP_NOALLOC_SCRATCH_BYTES)) +
             " of the static array; use dyng_hip_sssp_update with a workspace");
  }
  scoped_device guard(g->device);
  noalloc_device& st = noalloc_for(g->device, s);
  unsigned char* base = st.ptr;
  // Stamps, far marks and flags must be zero (or of an older generation) when a run starts, and a
  // run leaves them so; `ancestor` must hold no list tag (-2 - generation) of a later generation. Their place in the array moves with n: a call on another vertex count
  // (or after a run that stopped early) finds the last call's other arrays there, so it clears
  // them first.
  if (n != st.last_n || st.generation > INT32_MAX / 2) {
    clear_scratch(base + L.stamp, 4 * n, s);
    clear_scratch(base + L.in_far, 4 * n, s);
    clear_scratch(base + L.flag, 4 * n, s);
    clear_scratch(base + L.ancestor, 4 * n, s);
    st.generation = 0;
    st.last_n = n;
  }
  p.generation = ++st.generation;
  p.resident = 0;  // the old tree's words are always packed from its arrays
  attach_scratch(p, engine_scratch{reinterpret_cast<u64*>(base + L.packed), reinterpret_cast<int*>(base + L.stamp),
                                   reinterpret_cast<int*>(base + L.in_far), reinterpret_cast<int*>(base + L.flag),
                                   reinterpret_cast<int*>(base + L.ancestor), reinterpret_cast<int*>(base + L.cand),
                                   reinterpret_cast<int*>(base + L.frontier), reinterpret_cast<int*>(base + L.near_a),
                                   reinterpret_cast<int*>(base + L.near_b), reinterpret_cast<int*>(base + L.far_a),
                                   reinterpret_cast<int*>(base + L.far_b), reinterpret_cast<int*>(base + L.touched),
                                   reinterpret_cast<int4*>(base + L.chunks),
                                   L.chunk_cap, reinterpret_cast<control*>(base + L.ctl),
                                   reinterpret_cast<grid_sync_state*>(base + L.bar)});
  const int blocks = grid_blocks(run->tuning, st.occupancy, st.cu_count, p.from_scratch != 0);
  hipLaunchKernelGGL(k_reset_barrier, dim3(1), dim3(64), 0, s, p.bar, p.ctl);
  if (narrow_grid(blocks, st.occupancy_narrow, st.cu_count)) {
    hipLaunchKernelGGL(sssp_fused_kernel<4>, dim3(blocks), dim3(kBlock), 0, s, p);
  } else {
    hipLaunchKernelGGL(sssp_fused_kernel<8>, dim3(blocks), dim3(kBlock), 0, s, p);
  }
  DYNG_HIP_TRY(hipGetLastError());
  control* hc = &st.host_ctl;
  DYNG_HIP_TRY(hipMemcpyAsync(hc, p.ctl, sizeof(control), hipMemcpyDeviceToHost, s));
  DYNG_HIP_TRY(hipStreamSynchronize(s));
  if (hc->overflow != 0 || hc->error_bits != 0) {
    // a run that stopped early can leave flags, far marks and stamps of any later generation
    // behind: the next call clears them
    st.last_n = 0;  // the next call clears them
  }
  throw_on_errors(hc, timeout_us);
  st.generation = std::max(st.generation, hc->generation);
  read_counters(hc, true, p.pk.bits, blocks, counters);
  // no events (creating one may allocate): the kernel's own clock from its first to its last stamp
  counters.kernel_ms = hc->t_phase[7] >= hc->t_phase[0] ? (hc->t_phase[7] - hc->t_phase[0]) * 1e-5 : 0.0;
  if (out) *out = counters;
}

}  // namespace

}  // namespace dyng_hip

using namespace dyng_hip;

extern "C" int dyng_hip_workspace_create(int device, dyng_hip_workspace** out) {
  return guarded([&] {
    auto ws = std::make_unique<dyng_hip_workspace>();
    ws->device = device;
    *out = ws.release();
  });
}

extern "C" int dyng_hip_workspace_destroy(dyng_hip_workspace* ws) {
  return guarded([&] { delete ws; });
}

extern "C" int dyng_hip_sssp_update(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out) {
  return guarded([&] { run_engine(run, true, static_cast<hipStream_t>(stream), out); });
}

extern "C" int dyng_hip_sssp_compute(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out) {
  return guarded([&] { run_engine(run, false, static_cast<hipStream_t>(stream), out); });
}

extern "C" int dyng_hip_sssp_update_noalloc(const dyng_hip_sssp_run* run, void* stream, dyng_hip_counters* out) {
  return guarded([&] { run_engine_noalloc(run, static_cast<hipStream_t>(stream), out); });
}
