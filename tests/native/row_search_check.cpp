// Host check of the row searches of dyng_amd/csrc/row_search.hpp against std::lower_bound and
// std::upper_bound on the masked neighbour ids: random rows (runs of equal ids, dead marks in the
// sign bit) starting at every offset within a 16-byte word, every sub-range of short rows and
// random ranges of long ones, every probe value around the row's ids. Each row sits in an
// allocation of whole 16-byte words, as device allocations are, so a build with AddressSanitizer
// also shows that no load leaves the words that hold the row. Prints "ok <checks>" or the first
// mismatch and exits nonzero.
#define DYNG_ROW_SEARCH_FN inline
#include "../../dyng_amd/csrc/row_search.hpp"

#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

using dyng_hip::row_lower_bound;
using dyng_hip::row_upper_bound_from;

namespace {

int32_t id(int32_t x) { return x & 0x7fffffff; }

/// A row's entries at offset `skew` (0..3 entries) of an allocation of whole 16-byte words.
struct placed_row {
  std::vector<dyng_hip::row_quad> words;
  int skew = 0;
  int len = 0;
  placed_row(const std::vector<int32_t>& row, int s) : skew(s), len(static_cast<int>(row.size())) {
    words.resize((row.size() + s + 3) / 4 + 0);
    if (words.empty()) words.resize(1);
    for (auto& q : words) q.x[0] = q.x[1] = q.x[2] = q.x[3] = 0x7ffffff0;
    for (size_t i = 0; i < row.size(); ++i) data()[i] = row[i];
  }
  int32_t* data() { return &words[0].x[0] + skew; }
};

bool check_range(placed_row& pr, int lo, int hi, int32_t v, long long& checks) {
  const int32_t* a = pr.data();
  auto less_id = [](int32_t x, int32_t val) { return id(x) < val; };
  auto id_less = [](int32_t val, int32_t x) { return val < id(x); };
  const int want_lb = static_cast<int>(std::lower_bound(a + lo, a + hi, v, less_id) - a);
  const int want_ub = static_cast<int>(std::upper_bound(a + lo, a + hi, v, id_less) - a);
  const int lb = row_lower_bound(a, lo, hi, v);
  const int ub = row_upper_bound_from(a, lb, hi, v);
  ++checks;
  if (lb != want_lb || ub != want_ub) {
    std::printf("mismatch: len %d skew %d lo %d hi %d v %d: lower %d (want %d) upper %d (want %d)\n",
                pr.len, pr.skew, lo, hi, v, lb, want_lb, ub, want_ub);
    return false;
  }
  return true;
}

std::vector<int32_t> random_row(std::mt19937_64& rng, int len, int max_run) {
  std::vector<int32_t> row;
  int32_t x = static_cast<int32_t>(rng() % 5);
  while (static_cast<int>(row.size()) < len) {
    const int run = 1 + static_cast<int>(rng() % max_run);
    for (int r = 0; r < run && static_cast<int>(row.size()) < len; ++r) {
      row.push_back((rng() % 4 == 0) ? (x | static_cast<int32_t>(0x80000000u)) : x);
    }
    x += 1 + static_cast<int32_t>(rng() % 3);
  }
  return row;
}

}  // namespace

int main() {
  std::mt19937_64 rng(20251203);
  long long checks = 0;
  for (int len = 0; len <= 80; ++len) {
    for (int trial = 0; trial < 12; ++trial) {
      const auto row = random_row(rng, len, trial % 3 == 0 ? 40 : 3);
      placed_row pr(row, trial % 4);
      const int32_t top = row.empty() ? 3 : id(row.back()) + 3;
      for (int lo = 0; lo <= len; ++lo) {
        for (int hi = lo; hi <= len; hi += 1 + (hi - lo) / 8) {
          for (int32_t v = 0; v <= top; ++v) {
            if (!check_range(pr, lo, hi, v, checks)) return 1;
          }
        }
      }
    }
  }
  for (int len : {1000, 4097, 65536, 300001}) {
    for (int max_run : {1, 2, 50}) {
      const auto row = random_row(rng, len, max_run);
      placed_row pr(row, max_run % 4);
      const int32_t top = id(row.back()) + 2;
      for (int t = 0; t < 4000; ++t) {
        const int lo = static_cast<int>(rng() % (len + 1));
        const int hi = lo + static_cast<int>(rng() % (len - lo + 1));
        const int32_t v = static_cast<int32_t>(rng() % (top + 1));
        if (!check_range(pr, lo, hi, v, checks)) return 1;
        if (!check_range(pr, 0, len, v, checks)) return 1;
      }
    }
  }
  std::printf("ok %lld\n", checks);
  return 0;
}
