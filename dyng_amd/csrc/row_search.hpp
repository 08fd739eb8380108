// Searches in the sorted runs of a row (a neighbour id with the sign bit as the dead mark).
// Device code by default; a host build of the same functions (DYNG_ROW_SEARCH_FN defined as
// `inline`) checks them against std::lower_bound / std::upper_bound.
#pragma once

#include <cstdint>

#ifndef DYNG_ROW_SEARCH_FN
#define DYNG_ROW_SEARCH_FN __device__ __forceinline__
#endif

namespace dyng_hip {

/// Four entries of a row, 16-byte aligned: the last round of a search loads the aligned words
/// that hold its range with one instruction each.
struct alignas(16) row_quad {
  int32_t x[4];
};

/// First index in [lo, hi) whose neighbour (sign bit masked) is >= v, or hi. Each round probes
/// seven evenly spaced entries with independent loads, so a long row costs about log8 of its
/// length in dependent memory round trips instead of log2. The last 16 entries or fewer are read
/// as the (at most five) aligned 16-byte words that hold them, all in flight together: one more
/// round trip, and five load instructions where sixteen single entries would take sixteen (a
/// wave's lanes usually search different rows, so each instruction fetches up to 64 lines). A
/// word that holds an entry of [lo, hi) lies in the same page as that entry, so no load reaches
/// memory the row's allocation does not map.
DYNG_ROW_SEARCH_FN int row_lower_bound(const int32_t* a, int lo, int hi, int32_t v) {
  while (hi - lo > 16) {  // the answer lies in [lo, hi]
    const int step = (hi - lo) >> 3;
    int below = 0;  // probes lo + k * step (k = 1..7) whose neighbour is < v: a prefix of them
#pragma unroll
    for (int k = 1; k <= 7; ++k) below += (a[lo + k * step] & 0x7fffffff) < v ? 1 : 0;
    const int next_lo = below == 0 ? lo : lo + below * step + 1;
    hi = below == 7 ? hi : lo + (below + 1) * step;
    lo = next_lo;
  }
  if (lo >= hi) return lo;
  const int skew = static_cast<int>((reinterpret_cast<uintptr_t>(a + lo) & 15) >> 2);  // entries before a + lo
  const row_quad* q = reinterpret_cast<const row_quad*>(a + lo - skew);
  const int end = hi - lo + skew;  // one past the range, counted from q (at most 19)
  const int last = (end - 1) >> 2;  // the last word that holds an entry of the range
  row_quad w[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) w[k] = q[k < last ? k : last];  // no branch: all five in flight at once
  int below = 0;  // entries of [lo, hi) below v: a prefix of the range
#pragma unroll
  for (int k = 0; k < 5; ++k) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = 4 * k + c;
      below += (j >= skew && j < end && (w[k].x[c] & 0x7fffffff) < v) ? 1 : 0;
    }
  }
  return lo + below;
}

/// One past the last entry of v's run, given lb = row_lower_bound(a, ., hi, v). Parallel edges
/// are rare: the entries at lb and lb + 1 (loaded together) usually settle it; a longer run of v
/// is searched like a lower bound of v + 1.
DYNG_ROW_SEARCH_FN int row_upper_bound_from(const int32_t* a, int lb, int hi, int32_t v) {
  if (lb >= hi) return lb;
  const int32_t x0 = a[lb] & 0x7fffffff;
  const int32_t x1 = a[lb + 1 < hi ? lb + 1 : lb] & 0x7fffffff;
  if (x0 != v) return lb;
  if (lb + 1 >= hi || x1 != v) return lb + 1;
  return row_lower_bound(a, lb + 2, hi, v + 1);  // vertex ids are below INT32_MAX
}

}  // namespace dyng_hip
