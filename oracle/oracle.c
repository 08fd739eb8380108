/*
 * oracle.c -- CPU restatement of dynG's batch-dynamic SSSP path. TEST INFRASTRUCTURE ONLY: the
 * checker of the HIP product path and the CPU baseline of the benchmark (see oracle.h for what
 * pins each function). Paths in citations are relative to the reference root.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <limits.h>
#include <omp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;

#define PACKED_INF (~(u64)0)
/* output_max_distance = DISTANCE_INF / 2 - 1 (openmp.cpp:84). */
#define OUTPUT_MAX ((u64)(ORC_INF / 2 - 1))

static int fail(char* err, int errlen, int code, const char* fmt, ...) {
  if (err != NULL && errlen > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
  }
  return code;
}

static double now_s(void) { return omp_get_wtime(); }

void orc_free(void* p) { free(p); }

void orc_csr_free(orc_csr* g) {
  if (g == NULL) return;
  free(g->row_ptr);
  free(g->col);
  free(g->w);
  g->row_ptr = NULL;
  g->col = NULL;
  g->w = NULL;
}

void orc_applied_free(orc_applied* a) {
  if (a == NULL) return;
  orc_csr_free(&a->g);
  free(a->ins_src);
  free(a->ins_dst);
  free(a->weight_increased);
  free(a->del_src);
  free(a->del_dst);
  memset(a, 0, sizeof(*a));
}

/* ------------------------------------------------------------------------------------------------
 * Small growable int32 vector.
 * ----------------------------------------------------------------------------------------------*/
typedef struct ivec {
  int32_t* v;
  int64_t n, cap;
} ivec;

static void ivec_push(ivec* a, int32_t x) {
  if (a->n == a->cap) {
    a->cap = a->cap ? a->cap * 2 : 64;
    a->v = (int32_t*)realloc(a->v, (size_t)a->cap * sizeof(int32_t));
    if (a->v == NULL) abort();
  }
  a->v[a->n++] = x;
}

static void ivec_reserve(ivec* a, int64_t cap) {
  if (a->cap < cap) {
    a->cap = cap;
    a->v = (int32_t*)realloc(a->v, (size_t)cap * sizeof(int32_t));
    if (a->v == NULL && cap > 0) abort();
  }
}

static void ivec_free(ivec* a) {
  free(a->v);
  a->v = NULL;
  a->n = a->cap = 0;
}

/* ------------------------------------------------------------------------------------------------
 * Construction: build_rows / build_from_edges_host (cpp/src/graph/apply_host.cpp:527-690).
 * Counting sort by source keeps the input order inside a row; multi_edges::forbid merges repeated
 * (u,v) with the LAST occurrence's weights, keeping the first occurrence's position under
 * row_order::append; row_order::sorted sorts rows stably by neighbour id.
 * ----------------------------------------------------------------------------------------------*/
typedef struct build_item {
  int32_t vertex;
  int64_t first; /* position of the first occurrence (row order) */
  int64_t last;  /* input index whose weights are used */
} build_item;

static void stable_sort_items_by_vertex(build_item* a, int64_t n, build_item* tmp) {
  /* merge sort, stable */
  if (n < 2) return;
  int64_t h = n / 2;
  stable_sort_items_by_vertex(a, h, tmp);
  stable_sort_items_by_vertex(a + h, n - h, tmp);
  int64_t i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = (a[j].vertex < a[i].vertex) ? a[j++] : a[i++];
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, (size_t)n * sizeof(build_item));
}

static int cmp_item_first(const void* x, const void* y) {
  const build_item* a = (const build_item*)x;
  const build_item* b = (const build_item*)y;
  return (a->first > b->first) - (a->first < b->first);
}

int orc_build_from_edges(int64_t n, int32_t K, int64_t count, const int32_t* src,
                         const int32_t* dst, const int32_t* weights, const orc_props* props,
                         orc_csr* out, char* err, int errlen) {
  memset(out, 0, sizeof(*out));
  if (n < 0) return fail(err, errlen, ORC_INVALID, "num_vertices must be >= 0, got %lld", (long long)n);
  /* Undirected graphs store both arcs of every edge (a self-loop once). */
  int64_t total = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (src[i] < 0 || src[i] >= n || dst[i] < 0 || dst[i] >= n)
      return fail(err, errlen, ORC_INVALID, "edge %lld (%d, %d) is out of range [0, %lld)",
                  (long long)i, src[i], dst[i], (long long)n);
    if (src[i] == dst[i] && props->on_self_loop == 1) continue;
    if (src[i] == dst[i] && props->on_self_loop == 2)
      return fail(err, errlen, ORC_INVALID, "edge %lld is a self-loop", (long long)i);
    total += (!props->directed && src[i] != dst[i]) ? 2 : 1;
  }
  int32_t* s = (int32_t*)malloc((size_t)(total ? total : 1) * sizeof(int32_t));
  int32_t* d = (int32_t*)malloc((size_t)(total ? total : 1) * sizeof(int32_t));
  int64_t* idx = (int64_t*)malloc((size_t)(total ? total : 1) * sizeof(int64_t));
  int64_t t = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (src[i] == dst[i] && props->on_self_loop != 0) continue;
    s[t] = src[i];
    d[t] = dst[i];
    idx[t++] = i;
    if (!props->directed && src[i] != dst[i]) {
      s[t] = dst[i];
      d[t] = src[i];
      idx[t++] = i;
    }
  }
  int64_t* start = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
  for (int64_t i = 0; i < total; ++i) start[s[i] + 1]++;
  for (int64_t u = 0; u < n; ++u) start[u + 1] += start[u];
  int64_t* order = (int64_t*)malloc((size_t)(total ? total : 1) * sizeof(int64_t));
  {
    int64_t* cursor = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    memcpy(cursor, start, (size_t)(n + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < total; ++i) order[cursor[s[i]]++] = i;
    free(cursor);
  }
  build_item* kept = (build_item*)malloc((size_t)(total ? total : 1) * sizeof(build_item));
  build_item* row = (build_item*)malloc((size_t)(total ? total : 1) * sizeof(build_item));
  build_item* tmp = (build_item*)malloc((size_t)(total ? total : 1) * sizeof(build_item));
  int64_t* degree = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int64_t nk = 0;
  for (int64_t u = 0; u < n; ++u) {
    int64_t r = 0;
    for (int64_t j = start[u]; j < start[u + 1]; ++j) {
      row[r].vertex = d[order[j]];
      row[r].first = j;
      row[r].last = order[j];
      ++r;
    }
    if (!props->allow_parallel) {
      stable_sort_items_by_vertex(row, r, tmp);
      int64_t w = 0;
      for (int64_t q = 0; q < r; ++q) {
        if (w > 0 && row[w - 1].vertex == row[q].vertex) {
          row[w - 1].last = row[q].last;
        } else {
          row[w++] = row[q];
        }
      }
      r = w;
      if (!props->sorted_rows) qsort(row, (size_t)r, sizeof(build_item), cmp_item_first);
    } else if (props->sorted_rows) {
      stable_sort_items_by_vertex(row, r, tmp);
    }
    degree[u] = r;
    memcpy(kept + nk, row, (size_t)r * sizeof(build_item));
    nk += r;
  }
  out->n = n;
  out->m = nk;
  out->K = K;
  out->row_ptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  for (int64_t u = 0; u < n; ++u) out->row_ptr[u + 1] = out->row_ptr[u] + degree[u];
  out->col = (int32_t*)malloc((size_t)(nk ? nk : 1) * sizeof(int32_t));
  out->w = (int32_t*)malloc((size_t)((nk * K) > 0 ? nk * K : 1) * sizeof(int32_t));
  for (int64_t e = 0; e < nk; ++e) {
    out->col[e] = kept[e].vertex;
    for (int32_t k = 0; k < K; ++k) out->w[(int64_t)k * nk + e] = weights[idx[kept[e].last] * K + k];
  }
  free(s);
  free(d);
  free(idx);
  free(start);
  free(order);
  free(kept);
  free(row);
  free(tmp);
  free(degree);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------------------
 * Batch application: apply_batch_host (cpp/src/graph/apply_host.cpp:57-395), MOSP's
 * applyChangeBatch. Step 0 forms the effective operations (self-loops, vertex growth, undirected
 * expansion); changes are grouped by source row in batch order; every changed row is rebuilt as a
 * list of (neighbour, slot, alive) entries (deletions remove the first alive match, insertions
 * overwrite the first alive match or append); then the new rows are assembled and weight
 * increases classified by comparing the first (u,v) of row u before and after.
 * ----------------------------------------------------------------------------------------------*/
typedef struct row_entry {
  int32_t vertex;
  int64_t slot; /* >= 0: original edge; < 0: effective insertion -slot - 1 */
  int32_t alive;
} row_entry;

static void stable_sort_entries_by_vertex(row_entry* a, int64_t n, row_entry* tmp) {
  if (n < 2) return;
  int64_t h = n / 2;
  stable_sort_entries_by_vertex(a, h, tmp);
  stable_sort_entries_by_vertex(a + h, n - h, tmp);
  int64_t i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = (a[j].vertex < a[i].vertex) ? a[j++] : a[i++];
  while (i < h) tmp[k++] = a[i++];
  while (j < n) tmp[k++] = a[j++];
  memcpy(a, tmp, (size_t)n * sizeof(row_entry));
}

int orc_apply_batch(const orc_csr* g, const orc_props* props, int64_t num_inserts,
                    const int32_t* ins_src_in, const int32_t* ins_dst_in, const int32_t* ins_w,
                    int64_t num_deletes, const int32_t* del_src_in, const int32_t* del_dst_in,
                    int32_t threads, orc_applied* out, char* err, int errlen) {
  (void)threads;
  memset(out, 0, sizeof(*out));
  const int64_t n = g->n;
  const int64_t m = g->m;
  const int32_t K = g->K;
  orc_summary summary;
  memset(&summary, 0, sizeof(summary));

  /* --- Step 0 (apply_host.cpp:80-147) --- */
  int64_t cap_ins = 2 * num_inserts + 1, cap_del = 2 * num_deletes + 1;
  int32_t* ins_from = (int32_t*)malloc((size_t)cap_ins * sizeof(int32_t));
  int32_t* ins_to = (int32_t*)malloc((size_t)cap_ins * sizeof(int32_t));
  int64_t* ins_weights = (int64_t*)malloc((size_t)cap_ins * sizeof(int64_t));
  int32_t* del_from = (int32_t*)malloc((size_t)cap_del * sizeof(int32_t));
  int32_t* del_to = (int32_t*)malloc((size_t)cap_del * sizeof(int32_t));
  int64_t num_ins = 0, num_del = 0;
  int64_t n_after = n;
  int rc = ORC_OK;
  for (int64_t i = 0; i < num_inserts; ++i) {
    const int32_t u = ins_src_in[i], v = ins_dst_in[i];
    if (u < 0 || v < 0) {
      rc = fail(err, errlen, ORC_INVALID, "insertion %lld (%d, %d) has a negative id", (long long)i, u, v);
      goto done;
    }
    if (!props->allow_vertex_growth && (u >= n || v >= n)) {
      rc = fail(err, errlen, ORC_INVALID,
                "insertion %lld (%d, %d) names a vertex >= %lld and allow_vertex_growth is false",
                (long long)i, u, v, (long long)n);
      goto done;
    }
    if (u == v) {
      if (props->on_self_loop == 1) {
        ++summary.dropped_self_loops;
        continue;
      }
      if (props->on_self_loop == 2) {
        rc = fail(err, errlen, ORC_INVALID, "insertion %lld is a self-loop (%d, %d)", (long long)i, u, u);
        goto done;
      }
    }
    if ((u > v ? u : v) >= INT32_MAX) {
      rc = fail(err, errlen, ORC_INVALID, "insertion %lld names the largest representable vertex id", (long long)i);
      goto done;
    }
    int64_t hi = (int64_t)(u > v ? u : v) + 1;
    if (hi > n_after) n_after = hi;
    ins_from[num_ins] = u;
    ins_to[num_ins] = v;
    ins_weights[num_ins++] = i;
    if (!props->directed && u != v) {
      ins_from[num_ins] = v;
      ins_to[num_ins] = u;
      ins_weights[num_ins++] = i;
    }
  }
  for (int64_t j = 0; j < num_deletes; ++j) {
    const int32_t u = del_src_in[j], v = del_dst_in[j];
    if (u < 0 || v < 0) {
      rc = fail(err, errlen, ORC_INVALID, "deletion %lld (%d, %d) has a negative id", (long long)j, u, v);
      goto done;
    }
    if (u == v) {
      if (props->on_self_loop == 1) {
        ++summary.dropped_self_loops;
        continue;
      }
      if (props->on_self_loop == 2) {
        rc = fail(err, errlen, ORC_INVALID, "deletion %lld is a self-loop (%d, %d)", (long long)j, u, u);
        goto done;
      }
    }
    if (u >= n_after || v >= n_after) {
      int reps = (!props->directed && u != v) ? 2 : 1;
      for (int r = 0; r < reps; ++r) {
        if (props->on_missing_delete != 0) {
          rc = fail(err, errlen, ORC_INVALID, "deletion %lld (%d, %d) names a missing edge", (long long)j, u, v);
          goto done;
        }
        ++summary.ignored_deletions;
      }
      continue;
    }
    del_from[num_del] = u;
    del_to[num_del++] = v;
    if (!props->directed && u != v) {
      del_from[num_del] = v;
      del_to[num_del++] = u;
    }
  }
  {
    const int64_t rows = n_after;
    /* --- group by source row keeping batch order (apply_host.cpp:149-173) --- */
    int64_t* del_start = (int64_t*)calloc((size_t)rows + 2, sizeof(int64_t));
    int64_t* ins_start = (int64_t*)calloc((size_t)rows + 2, sizeof(int64_t));
    for (int64_t j = 0; j < num_del; ++j) del_start[del_from[j] + 1]++;
    for (int64_t i = 0; i < num_ins; ++i) ins_start[ins_from[i] + 1]++;
    for (int64_t u = 0; u < rows; ++u) {
      del_start[u + 1] += del_start[u];
      ins_start[u + 1] += ins_start[u];
    }
    int64_t* del_order = (int64_t*)malloc((size_t)(num_del + 1) * sizeof(int64_t));
    int64_t* ins_order = (int64_t*)malloc((size_t)(num_ins + 1) * sizeof(int64_t));
    int64_t* cursor = (int64_t*)malloc((size_t)(rows + 1) * sizeof(int64_t));
    memcpy(cursor, del_start, (size_t)rows * sizeof(int64_t));
    for (int64_t j = 0; j < num_del; ++j) del_order[cursor[del_from[j]]++] = j;
    memcpy(cursor, ins_start, (size_t)rows * sizeof(int64_t));
    for (int64_t i = 0; i < num_ins; ++i) ins_order[cursor[ins_from[i]]++] = i;
    free(cursor);

    /* --- rebuild every changed row (apply_host.cpp:175-252) --- */
    int64_t* row_index = (int64_t*)malloc((size_t)(rows + 1) * sizeof(int64_t));
    for (int64_t u = 0; u < rows; ++u) row_index[u] = -1;
    int64_t nrows = 0;
    row_entry** contents = NULL;
    int64_t* content_len = NULL;
    int64_t contents_cap = 0;
    for (int64_t u = 0; u < rows && rc == ORC_OK; ++u) {
      if (del_start[u] == del_start[u + 1] && ins_start[u] == ins_start[u + 1]) continue;
      int64_t old_deg = u < n ? g->row_ptr[u + 1] - g->row_ptr[u] : 0;
      int64_t cap = old_deg + (ins_start[u + 1] - ins_start[u]) + 1;
      row_entry* row = (row_entry*)malloc((size_t)cap * sizeof(row_entry));
      int64_t len = 0;
      if (u < n) {
        for (int64_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e) {
          row[len].vertex = g->col[e];
          row[len].slot = e;
          row[len].alive = 1;
          ++len;
        }
      }
      for (int phase = 0; phase < 2 && rc == ORC_OK; ++phase) {
        const int deletions = props->deletions_first ? (phase == 0) : (phase == 1);
        if (deletions) {
          for (int64_t j = del_start[u]; j < del_start[u + 1]; ++j) {
            const int32_t v = del_to[del_order[j]];
            int removed = 0;
            for (int64_t q = 0; q < len; ++q) {
              if (row[q].alive && row[q].vertex == v) {
                row[q].alive = 0;
                removed = 1;
                break;
              }
            }
            if (removed) {
              ++summary.deleted_edges;
            } else {
              if (props->on_missing_delete != 0) {
                rc = fail(err, errlen, ORC_INVALID, "deletion (%lld, %d) names a missing edge", (long long)u, v);
                break;
              }
              ++summary.ignored_deletions;
            }
          }
        } else {
          for (int64_t j = ins_start[u]; j < ins_start[u + 1]; ++j) {
            const int64_t i = ins_order[j];
            const int32_t v = ins_to[i];
            int replaced = 0;
            for (int64_t q = 0; q < len; ++q) {
              if (row[q].alive && row[q].vertex == v) {
                if (props->on_existing_insert == 0) {
                  row[q].slot = -i - 1;
                  ++summary.updated_edges;
                } else if (props->on_existing_insert == 2) {
                  ++summary.ignored_insertions;
                } else {
                  rc = fail(err, errlen, ORC_INVALID, "insertion %lld (%lld, %d) names an existing edge",
                            (long long)ins_weights[i], (long long)u, v);
                }
                replaced = 1;
                break;
              }
            }
            if (rc != ORC_OK) break;
            if (!replaced) {
              row[len].vertex = v;
              row[len].slot = -i - 1;
              row[len].alive = 1;
              ++len;
              ++summary.inserted_edges;
            }
          }
        }
      }
      if (props->sorted_rows) {
        int64_t w = 0;
        for (int64_t q = 0; q < len; ++q)
          if (row[q].alive) row[w++] = row[q];
        len = w;
        row_entry* tmp = (row_entry*)malloc((size_t)(len + 1) * sizeof(row_entry));
        stable_sort_entries_by_vertex(row, len, tmp);
        free(tmp);
      }
      if (nrows == contents_cap) {
        contents_cap = contents_cap ? 2 * contents_cap : 64;
        contents = (row_entry**)realloc(contents, (size_t)contents_cap * sizeof(row_entry*));
        content_len = (int64_t*)realloc(content_len, (size_t)contents_cap * sizeof(int64_t));
      }
      row_index[u] = nrows;
      contents[nrows] = row;
      content_len[nrows] = len;
      ++nrows;
    }
    if (rc == ORC_OK) {
      /* --- assemble (apply_host.cpp:254-346) --- */
      out->g.n = rows;
      out->g.K = K;
      out->g.row_ptr = (int64_t*)calloc((size_t)rows + 1, sizeof(int64_t));
      for (int64_t u = 0; u < rows; ++u) {
        int64_t dg = 0;
        if (row_index[u] >= 0) {
          for (int64_t q = 0; q < content_len[row_index[u]]; ++q) dg += contents[row_index[u]][q].alive ? 1 : 0;
        } else if (u < n) {
          dg = g->row_ptr[u + 1] - g->row_ptr[u];
        }
        out->g.row_ptr[u + 1] = out->g.row_ptr[u] + dg;
      }
      const int64_t m_new = out->g.row_ptr[rows];
      out->g.m = m_new;
      out->g.col = (int32_t*)malloc((size_t)(m_new ? m_new : 1) * sizeof(int32_t));
      out->g.w = (int32_t*)malloc((size_t)((m_new * K) > 0 ? m_new * K : 1) * sizeof(int32_t));
      for (int64_t u = 0; u < rows; ++u) {
        int64_t o = out->g.row_ptr[u];
        if (row_index[u] < 0) {
          if (u >= n) continue;
          for (int64_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e, ++o) {
            out->g.col[o] = g->col[e];
            for (int32_t k = 0; k < K; ++k) out->g.w[(int64_t)k * m_new + o] = g->w[(int64_t)k * m + e];
          }
          continue;
        }
        const row_entry* row = contents[row_index[u]];
        for (int64_t q = 0; q < content_len[row_index[u]]; ++q) {
          if (!row[q].alive) continue;
          out->g.col[o] = row[q].vertex;
          if (row[q].slot >= 0) {
            for (int32_t k = 0; k < K; ++k) out->g.w[(int64_t)k * m_new + o] = g->w[(int64_t)k * m + row[q].slot];
          } else {
            const int64_t i = ins_weights[-row[q].slot - 1];
            for (int32_t k = 0; k < K; ++k) out->g.w[(int64_t)k * m_new + o] = ins_w[i * K + k];
          }
          ++o;
        }
      }
      summary.inserted_vertices = n_after - n;
      summary.num_vertices_after = n_after;
      /* --- weight-increase classification (apply_host.cpp:348-393) --- */
      out->num_ins = num_ins;
      out->ins_src = (int32_t*)malloc((size_t)(num_ins + 1) * sizeof(int32_t));
      out->ins_dst = (int32_t*)malloc((size_t)(num_ins + 1) * sizeof(int32_t));
      memcpy(out->ins_src, ins_from, (size_t)num_ins * sizeof(int32_t));
      memcpy(out->ins_dst, ins_to, (size_t)num_ins * sizeof(int32_t));
      out->weight_increased = (uint8_t*)calloc((size_t)(num_ins * K + 1), 1);
      out->num_del = num_del;
      out->del_src = (int32_t*)malloc((size_t)(num_del + 1) * sizeof(int32_t));
      out->del_dst = (int32_t*)malloc((size_t)(num_del + 1) * sizeof(int32_t));
      memcpy(out->del_src, del_from, (size_t)num_del * sizeof(int32_t));
      memcpy(out->del_dst, del_to, (size_t)num_del * sizeof(int32_t));
      for (int64_t i = 0; i < num_ins; ++i) {
        const int64_t u = ins_from[i];
        const int32_t v = ins_to[i];
        if (u >= n) continue;
        int64_t original_edge = -1, final_edge = -1;
        for (int64_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e)
          if (g->col[e] == v) {
            original_edge = e;
            break;
          }
        if (original_edge < 0) continue;
        for (int64_t e = out->g.row_ptr[u]; e < out->g.row_ptr[u + 1]; ++e)
          if (out->g.col[e] == v) {
            final_edge = e;
            break;
          }
        if (final_edge < 0) continue;
        for (int32_t k = 0; k < K; ++k)
          if (out->g.w[(int64_t)k * m_new + final_edge] > g->w[(int64_t)k * m + original_edge])
            out->weight_increased[i * K + k] = 1;
      }
      out->summary = summary;
    }
    for (int64_t r = 0; r < nrows; ++r) free(contents[r]);
    free(contents);
    free(content_len);
    free(row_index);
    free(del_start);
    free(ins_start);
    free(del_order);
    free(ins_order);
  }
done:
  free(ins_from);
  free(ins_to);
  free(ins_weights);
  free(del_from);
  free(del_to);
  if (rc != ORC_OK) orc_applied_free(out);
  return rc;
}

/* ------------------------------------------------------------------------------------------------
 * transpose_host (cpp/src/graph/apply_host.cpp:474-518): in-rows list their sources in out-edge
 * order.
 * ----------------------------------------------------------------------------------------------*/
int orc_transpose(const orc_csr* g, orc_csr* out, int32_t threads) {
  (void)threads;
  const int64_t n = g->n, m = g->m;
  const int32_t K = g->K;
  memset(out, 0, sizeof(*out));
  out->n = n;
  out->m = m;
  out->K = K;
  out->row_ptr = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
  for (int64_t e = 0; e < m; ++e) out->row_ptr[g->col[e] + 1]++;
  for (int64_t v = 0; v < n; ++v) out->row_ptr[v + 1] += out->row_ptr[v];
  out->col = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
  out->w = (int32_t*)malloc((size_t)((m * K) > 0 ? m * K : 1) * sizeof(int32_t));
  int64_t* cursor = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
  memcpy(cursor, out->row_ptr, (size_t)(n + 1) * sizeof(int64_t));
  int64_t* position = (int64_t*)malloc((size_t)(m ? m : 1) * sizeof(int64_t));
  for (int64_t u = 0; u < n; ++u) {
    for (int64_t e = g->row_ptr[u]; e < g->row_ptr[u + 1]; ++e) {
      const int64_t pos = cursor[g->col[e]]++;
      position[e] = pos;
      out->col[pos] = (int32_t)u;
    }
  }
  for (int32_t k = 0; k < K; ++k)
    for (int64_t e = 0; e < m; ++e) out->w[(int64_t)k * m + position[e]] = g->w[(int64_t)k * m + e];
  free(cursor);
  free(position);
  return ORC_OK;
}

/* build_changes (cpp/src/algorithms/sssp/sssp.cpp:563-576): every deletion, then every insertion
 * that raised objective k's weight. */
int64_t orc_build_changes(const orc_applied* a, int32_t k, int32_t** from, int32_t** to) {
  const int32_t K = a->g.K;
  int64_t count = a->num_del;
  for (int64_t i = 0; i < a->num_ins; ++i) count += a->weight_increased[i * K + k] ? 1 : 0;
  *from = (int32_t*)malloc((size_t)(count + 1) * sizeof(int32_t));
  *to = (int32_t*)malloc((size_t)(count + 1) * sizeof(int32_t));
  int64_t c = 0;
  for (int64_t j = 0; j < a->num_del; ++j) {
    (*from)[c] = a->del_src[j];
    (*to)[c++] = a->del_dst[j];
  }
  for (int64_t i = 0; i < a->num_ins; ++i) {
    if (a->weight_increased[i * K + k]) {
      (*from)[c] = a->ins_src[i];
      (*to)[c++] = a->ins_dst[i];
    }
  }
  return count;
}

/* ------------------------------------------------------------------------------------------------
 * Closed forms of sssp.cpp:136-165 and the device packing rule (kernels.cuh:82-92).
 * ----------------------------------------------------------------------------------------------*/
int64_t orc_default_delta(int64_t m, int64_t n, int64_t weight_sum) {
  if (m <= 0 || n <= 0) return 1;
  const double average_weight = (double)weight_sum / (double)m;
  const double average_degree = (double)m / (double)n;
  const int64_t d = (int64_t)(32.0 * average_weight / average_degree);
  return d > 1 ? d : 1;
}

static int parent_bits_for(int64_t n) {
  int bits = 1;
  while ((((u64)1) << bits) - 1 < (u64)n) ++bits;
  return bits;
}

int32_t orc_distances_fit(int64_t n, int64_t max_weight) {
  const u64 weight = (u64)(max_weight > 1 ? max_weight : 1);
  const u64 hops = (u64)(n - 1 > 1 ? n - 1 : 1);
  return weight <= OUTPUT_MAX / hops;
}

int32_t orc_packs_parents(int64_t n, int64_t max_weight) {
  const u64 weight = (u64)(max_weight > 1 ? max_weight : 1);
  const u64 hops = (u64)(n - 1 > 1 ? n - 1 : 1);
  const int bits = parent_bits_for(n);
  const u64 max_distance = (PACKED_INF >> bits) - 1;
  const u64 largest_candidate = weight * hops + weight;
  return weight <= max_distance / hops && largest_candidate <= max_distance;
}

int32_t orc_device_parent_bits(int64_t n, int64_t max_weight) {
  const u64 weight = (u64)(max_weight > 1 ? max_weight : 1);
  const u64 hops = (u64)(n - 1 > 1 ? n - 1 : 1);
  const int bits = parent_bits_for(n);
  const u64 max_distance = (PACKED_INF >> bits) - 1;
  return weight * hops <= max_distance ? bits : 0;
}

/* ------------------------------------------------------------------------------------------------
 * Workspace (problem.hpp:144-205): arrays sized once and reused; stamps deduplicate list entries
 * by generation; in_far is zero between runs.
 * ----------------------------------------------------------------------------------------------*/
struct orc_ws {
  int64_t capacity;
  int generation;
  u64* packed;
  int* stamp;
  char* in_far;
  char* state;
  int32_t* saved_parents;
  ivec near_a, near_b, far, far2, candidates, frontier;
  /* sequential engine */
  int32_t* child_start;
  ivec children, invalid, affected, candidate_list, touched, old_parent;
  int64_t* old_distance;
  int64_t old_distance_cap;
  /* per-thread lists: 2 per thread */
  ivec* tl;
  int tl_count;
};

orc_ws* orc_ws_new(void) { return (orc_ws*)calloc(1, sizeof(orc_ws)); }

void orc_ws_free(orc_ws* ws) {
  if (ws == NULL) return;
  free(ws->packed);
  free(ws->stamp);
  free(ws->in_far);
  free(ws->state);
  free(ws->saved_parents);
  free(ws->child_start);
  free(ws->old_distance);
  ivec* lists[] = {&ws->near_a, &ws->near_b, &ws->far, &ws->far2, &ws->candidates, &ws->frontier,
                   &ws->children, &ws->invalid, &ws->affected, &ws->candidate_list, &ws->touched,
                   &ws->old_parent};
  for (size_t i = 0; i < sizeof(lists) / sizeof(lists[0]); ++i) ivec_free(lists[i]);
  for (int t = 0; t < ws->tl_count; ++t) ivec_free(&ws->tl[t]);
  free(ws->tl);
  free(ws);
}

static void ws_reserve(orc_ws* ws, int64_t n, int threads) {
  if (n > ws->capacity) {
    const size_t c = (size_t)(n > 0 ? n : 1);
    free(ws->packed);
    free(ws->stamp);
    free(ws->in_far);
    free(ws->state);
    free(ws->child_start);
    free(ws->saved_parents);
    ws->packed = (u64*)malloc(c * sizeof(u64));
    for (size_t i = 0; i < c; ++i) ws->packed[i] = PACKED_INF;
    ws->stamp = (int*)calloc(c, sizeof(int));
    ws->in_far = (char*)calloc(c, 1);
    ws->state = (char*)calloc(c, 1);
    ws->child_start = (int32_t*)calloc(c + 1, sizeof(int32_t));
    ws->saved_parents = (int32_t*)malloc(c * sizeof(int32_t));
    ivec* lists[] = {&ws->near_a, &ws->near_b, &ws->far, &ws->far2, &ws->candidates, &ws->frontier};
    for (size_t i = 0; i < 6; ++i) {
      lists[i]->n = 0;
      ivec_reserve(lists[i], (int64_t)c);
    }
    ws->capacity = n;
    ws->generation = 0;
  }
  if (ws->tl_count < 2 * threads) {
    ws->tl = (ivec*)realloc(ws->tl, (size_t)(2 * threads) * sizeof(ivec));
    for (int t = ws->tl_count; t < 2 * threads; ++t) memset(&ws->tl[t], 0, sizeof(ivec));
    ws->tl_count = 2 * threads;
  }
}

static int next_generation(orc_ws* ws) {
  if (ws->generation == INT_MAX) {
    memset(ws->stamp, 0, (size_t)ws->capacity * sizeof(int));
    ws->generation = 0;
  }
  return ++ws->generation;
}

/* ================================================================================================
 * The sequential engine: sequential.cpp:57-381 (the reference backend).
 * ==============================================================================================*/
typedef struct seq_engine {
  orc_run* run;
  orc_ws* ws;
  int64_t n;
  int affected_generation;
  int count_changes;
  int packed_parents;
} seq_engine;

static void seq_save(seq_engine* e, int32_t v) {
  if (!e->count_changes) return;
  orc_ws* ws = e->ws;
  if (ws->state[v] != 0) return;
  ws->state[v] = 1;
  ivec_push(&ws->touched, v);
  if (ws->touched.n > ws->old_distance_cap) {
    ws->old_distance_cap = ws->touched.cap;
    ws->old_distance = (int64_t*)realloc(ws->old_distance, (size_t)ws->old_distance_cap * sizeof(int64_t));
  }
  ws->old_distance[ws->touched.n - 1] = e->run->dist[v];
  ivec_push(&ws->old_parent, e->run->parent[v]);
}

static void seq_init(seq_engine* e, orc_run* run, orc_ws* ws) {
  e->run = run;
  e->ws = ws;
  e->n = run->n;
  e->count_changes = 1;
  ws_reserve(ws, run->n, 1);
  run->invalidated = run->iterations = run->epochs = run->pushes = run->affected = 0;
  e->packed_parents = orc_packs_parents(run->n, run->max_weight);
  run->packed_parents = e->packed_parents;
  ws->touched.n = 0;
  ws->old_parent.n = 0;
  memset(ws->state, 0, (size_t)(run->n > 0 ? run->n : 0));
}

/* relax (sequential.cpp:349-381): re-evaluate v over its in-neighbours. */
static int seq_relax(seq_engine* e, int32_t v) {
  orc_run* r = e->run;
  int32_t best_parent = -1;
  int64_t best_distance = ORC_INF;
  for (int64_t q = r->in_row[v]; q < r->in_row[v + 1]; ++q) {
    const int32_t u = r->in_col[q];
    if (r->dist[u] >= ORC_INF / 2) continue;
    const int64_t cd = r->dist[u] + (int64_t)r->in_w[q];
    if (cd < best_distance || (cd == best_distance && u < best_parent)) {
      best_distance = cd;
      best_parent = u;
    }
  }
  const int better = best_distance < r->dist[v] ||
                     (best_distance == r->dist[v] && best_parent >= 0 && best_parent < r->parent[v]);
  if (!better) return 0;
  const int decreased = best_distance < r->dist[v];
  seq_save(e, v);
  r->dist[v] = best_distance;
  r->parent[v] = best_parent;
  return decreased;
}

static void seq_mark_affected(seq_engine* e, int32_t v) {
  if (e->ws->stamp[v] != e->affected_generation) {
    e->ws->stamp[v] = e->affected_generation;
    ivec_push(&e->ws->affected, v);
  }
}

/* expect_no_rootless_cycle (sequential.cpp:293-320). */
static int seq_no_rootless_cycle(seq_engine* e, int invalid_generation) {
  const int32_t* parent = e->run->parent;
  char* color = e->ws->state;
  ivec* path = &e->ws->affected;
  int cyclic = 0;
  for (int64_t v = 0; v < e->n && !cyclic; ++v) {
    path->n = 0;
    int32_t u = (int32_t)v;
    while (u >= 0 && color[u] == 0 && e->ws->stamp[u] != invalid_generation) {
      color[u] = 1;
      ivec_push(path, u);
      u = parent[u];
    }
    cyclic = u >= 0 && color[u] == 1;
    for (int64_t i = 0; i < path->n; ++i) color[path->v[i]] = 2;
  }
  memset(color, 0, (size_t)e->ws->capacity);
  path->n = 0;
  return !cyclic;
}

static int seq_identify_affected(seq_engine* e, char* err, int errlen) {
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  const int invalid_generation = next_generation(ws);
  ivec* inv = &ws->invalid;
  inv->n = 0;
  for (int64_t i = 0; i < r->num_changed; ++i) {
    const int32_t u = r->changed_from[i], v = r->changed_to[i];
    if (r->parent[v] == u && ws->stamp[v] != invalid_generation) {
      ws->stamp[v] = invalid_generation;
      ivec_push(inv, v);
    }
  }
  if (r->num_changed > 0 && !seq_no_rootless_cycle(e, invalid_generation))
    return fail(err, errlen, ORC_INVALID, "sssp: the input shortest-path tree has a parent cycle");
  if (inv->n > 0) {
    int32_t* cs = ws->child_start;
    memset(cs, 0, (size_t)(e->n + 1) * sizeof(int32_t));
    for (int64_t v = 0; v < e->n; ++v)
      if (v != r->source && r->parent[v] >= 0) cs[r->parent[v] + 1]++;
    for (int64_t v = 0; v < e->n; ++v) cs[v + 1] += cs[v];
    ivec_reserve(&ws->children, cs[e->n] + 1);
    ws->children.n = cs[e->n];
    ivec* cursor = &ws->candidate_list;
    ivec_reserve(cursor, e->n + 1);
    memcpy(cursor->v, cs, (size_t)e->n * sizeof(int32_t));
    for (int64_t v = 0; v < e->n; ++v)
      if (v != r->source && r->parent[v] >= 0) ws->children.v[cursor->v[r->parent[v]]++] = (int32_t)v;
    for (int64_t i = 0; i < inv->n; ++i) {
      const int32_t x = inv->v[i];
      for (int32_t c = cs[x]; c < cs[x + 1]; ++c) {
        const int32_t child = ws->children.v[c];
        if (ws->stamp[child] != invalid_generation) {
          ws->stamp[child] = invalid_generation;
          ivec_push(inv, child);
        }
      }
    }
    for (int64_t i = 0; i < inv->n; ++i) {
      const int32_t v = inv->v[i];
      seq_save(e, v);
      r->dist[v] = ORC_INF;
      r->parent[v] = -1;
    }
  }
  r->invalidated = inv->n;
  return ORC_OK;
}

static void seq_seed(seq_engine* e) {
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  ws->affected.n = 0;
  e->affected_generation = next_generation(ws);
  for (int64_t i = 0; i < ws->invalid.n; ++i)
    if (seq_relax(e, ws->invalid.v[i])) seq_mark_affected(e, ws->invalid.v[i]);
  for (int64_t i = 0; i < r->num_insert_heads; ++i) {
    const int32_t v = r->insert_heads[i];
    if (v != r->source && seq_relax(e, v)) seq_mark_affected(e, v);
  }
}

static int seq_loop(seq_engine* e, char* err, int errlen) {
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  ivec* next = &ws->candidate_list;
  while (ws->affected.n > 0) {
    ++r->iterations;
    if (r->iterations > e->n) return fail(err, errlen, ORC_INTERNAL, "sssp: the sequential update did not converge");
    r->pushes += ws->affected.n;
    const int next_gen = next_generation(ws);
    next->n = 0;
    for (int64_t i = 0; i < ws->affected.n; ++i) {
      const int32_t a = ws->affected.v[i];
      const int64_t da = r->dist[a];
      if (da >= ORC_INF / 2) continue;
      for (int64_t q = r->out_row[a]; q < r->out_row[a + 1]; ++q) {
        const int32_t x = r->out_col[q];
        if (x == r->source) continue;
        const int64_t offer = da + (int64_t)r->out_w[q];
        if (offer < r->dist[x] || (offer == r->dist[x] && a < r->parent[x])) {
          const int decreased = offer < r->dist[x];
          seq_save(e, x);
          r->dist[x] = offer;
          r->parent[x] = a;
          if (decreased && ws->stamp[x] != next_gen) {
            ws->stamp[x] = next_gen;
            ivec_push(next, x);
          }
        }
      }
    }
    ivec t = ws->affected;
    ws->affected = *next;
    *next = t;
  }
  return ORC_OK;
}

static void seq_finalize(seq_engine* e) {
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  if (!e->packed_parents) {
    /* recover_parents (sequential.cpp:247-268) */
    for (int64_t v = 0; v < e->n; ++v) {
      int32_t best = -1;
      if (v != r->source && r->dist[v] < ORC_INF / 2) {
        for (int64_t q = r->in_row[v]; q < r->in_row[v + 1]; ++q) {
          const int32_t u = r->in_col[q];
          if (r->dist[u] < ORC_INF / 2 && r->dist[u] + (int64_t)r->in_w[q] == r->dist[v] && (best < 0 || u < best))
            best = u;
        }
      }
      if (best != r->parent[v]) {
        seq_save(e, (int32_t)v);
        r->parent[v] = best;
      }
    }
  }
  /* count_affected (sequential.cpp:271-288) */
  if (!e->count_changes) {
    int64_t reachable = 0;
    for (int64_t v = 0; v < e->n; ++v) reachable += r->dist[v] < ORC_INF / 2 ? 1 : 0;
    r->affected = reachable;
    return;
  }
  int64_t affected = 0;
  for (int64_t i = 0; i < ws->touched.n; ++i) {
    const int32_t v = ws->touched.v[i];
    affected += (r->dist[v] != ws->old_distance[i] || r->parent[v] != ws->old_parent.v[i]) ? 1 : 0;
  }
  r->affected = affected;
}

int orc_seq_compute(orc_run* run, orc_ws* ws_in, char* err, int errlen) {
  orc_ws* ws = ws_in ? ws_in : orc_ws_new();
  seq_engine e;
  seq_init(&e, run, ws);
  int rc = ORC_OK;
  if (run->source < 0 || run->source >= run->n) {
    rc = fail(err, errlen, ORC_INVALID, "sssp::compute: source %d is out of range [0, %lld)", run->source, (long long)run->n);
  } else {
    /* reset + seed_static (sequential.cpp:84-97) */
    for (int64_t v = 0; v < run->n; ++v) {
      run->dist[v] = ORC_INF;
      run->parent[v] = -1;
    }
    run->dist[run->source] = 0;
    e.count_changes = 0;
    ws->affected.n = 0;
    ivec_push(&ws->affected, run->source);
    double t0 = now_s();
    rc = seq_loop(&e, err, errlen);
    double t1 = now_s();
    if (rc == ORC_OK) seq_finalize(&e);
    run->t_hooks[0] = 0;
    run->t_hooks[1] = 0;
    run->t_hooks[2] = t1 - t0;
    run->t_hooks[3] = now_s() - t1;
  }
  if (!ws_in) orc_ws_free(ws);
  return rc;
}

int orc_seq_update(orc_run* run, orc_ws* ws_in, char* err, int errlen) {
  orc_ws* ws = ws_in ? ws_in : orc_ws_new();
  seq_engine e;
  seq_init(&e, run, ws);
  double t0 = now_s();
  int rc = seq_identify_affected(&e, err, errlen);
  double t1 = now_s(), t2 = t1, t3 = t1;
  if (rc == ORC_OK) {
    seq_seed(&e);
    t2 = now_s();
    rc = seq_loop(&e, err, errlen);
    t3 = now_s();
  }
  if (rc == ORC_OK) seq_finalize(&e);
  run->t_hooks[0] = t1 - t0;
  run->t_hooks[1] = t2 - t1;
  run->t_hooks[2] = t3 - t2;
  run->t_hooks[3] = now_s() - t3;
  if (!ws_in) orc_ws_free(ws);
  return rc;
}

/* ================================================================================================
 * The OpenMP engine: openmp.cpp:92-583 (MOSP-OpenMP's sospUpdateCpu), the CPU baseline.
 * Per-thread lists are gathered into the shared list after each work-sharing loop
 * (util/list_gather.hpp); list order does not change any result (packed atomic minima).
 * ==============================================================================================*/
typedef struct packing {
  int bits;
  u64 no_parent;
} packing;

static inline u64 pk_pack(packing p, u64 d, int32_t parent) {
  if (p.bits == 0) return d;
  return (d << p.bits) | (parent < 0 ? p.no_parent : (u64)(uint32_t)parent);
}
static inline u64 pk_dist(packing p, u64 w) { return w >> p.bits; }
static inline int32_t pk_parent(packing p, u64 w) {
  const u64 q = w & p.no_parent;
  return q == p.no_parent ? -1 : (int32_t)q;
}

static inline u64 load64(const u64* p) { return __atomic_load_n(p, __ATOMIC_RELAXED); }

static inline u64 atomic_min64(u64* p, u64 value) {
  u64 old = load64(p);
  while (value < old && !__atomic_compare_exchange_n(p, &old, value, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
  return old;
}

static inline void atomic_min_id(int32_t* p, int32_t value) {
  int32_t old = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (value < old && !__atomic_compare_exchange_n(p, &old, value, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
  }
}

static inline int claim(int* stamp, int32_t v, int generation) {
  return __atomic_exchange_n(&stamp[v], generation, __ATOMIC_RELAXED) != generation;
}

/* Gather per-thread lists into `out` (inside a parallel region; every thread calls it). */
static void gather(ivec* out, ivec* tl, int slot, int64_t* offsets) {
  const int t = omp_get_thread_num();
  const int T = omp_get_num_threads();
#pragma omp barrier
#pragma omp single
  {
    int64_t total = out->n;
    for (int i = 0; i < T; ++i) {
      offsets[i] = total;
      total += tl[2 * i + slot].n;
    }
    ivec_reserve(out, total + 1);
    out->n = total;
  }
  memcpy(out->v + offsets[t], tl[2 * t + slot].v, (size_t)tl[2 * t + slot].n * sizeof(int32_t));
#pragma omp barrier
}

typedef struct omp_engine {
  orc_run* run;
  orc_ws* ws;
  int threads;
  packing pk;
  u64 bound;
  int empty;
  int64_t offsets[2][1024];
} omp_engine;

static int omp_init(omp_engine* e, orc_run* run, orc_ws* ws, char* err, int errlen) {
  e->run = run;
  e->ws = ws;
  e->threads = run->threads > 0 ? run->threads : 1;
  if (e->threads > 1024) e->threads = 1024;
  run->invalidated = run->iterations = run->epochs = run->pushes = run->affected = 0;
  run->packed_parents = 1;
  e->empty = run->n == 0;
  if (e->empty) return ORC_OK;
  if (run->delta <= 0) return fail(err, errlen, ORC_INVALID, "sssp: the near-far width delta must be > 0, got %lld", (long long)run->delta);
  ws_reserve(ws, run->n, e->threads);
  /* choose_packing (openmp.cpp:147-158) */
  const u64 weight = (u64)(run->max_weight > 1 ? run->max_weight : 1);
  const u64 hops = (u64)(run->n - 1 > 1 ? run->n - 1 : 1);
  if (weight > OUTPUT_MAX / hops) return fail(err, errlen, ORC_INVALID, "sssp: distances do not fit in 62 bits");
  e->bound = weight * hops;
  const int bits = parent_bits_for(run->n);
  const u64 max_distance = (PACKED_INF >> bits) - 1;
  if (e->bound + weight <= max_distance) {
    e->pk.bits = bits;
    e->pk.no_parent = (((u64)1) << bits) - 1;
  } else {
    e->pk.bits = 0;
    e->pk.no_parent = 0;
  }
  run->packed_parents = e->pk.bits > 0;
  return ORC_OK;
}

static int omp_identify_affected(omp_engine* e, char* err, int errlen) {
  if (e->empty) return ORC_OK;
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  const int64_t n = r->n;
  u64* packed = ws->packed;
  char* state = ws->state;
  int* stamp = ws->stamp;
  const packing pk = e->pk;
  const u64 bound = e->bound;
  int overflow = 0;
#pragma omp parallel for num_threads(e->threads) schedule(static) reduction(|| : overflow)
  for (int64_t v = 0; v < n; ++v) {
    const int64_t d = r->dist[v];
    state[v] = 0;
    if (d >= ORC_INF / 2) {
      packed[v] = PACKED_INF;
    } else if (d < 0 || (u64)d > bound) {
      overflow = 1;
      packed[v] = PACKED_INF;
    } else {
      packed[v] = pk_pack(pk, (u64)d, r->parent[v]);
    }
  }
  if (overflow) return fail(err, errlen, ORC_INVALID, "sssp: an input distance exceeds (n - 1) * max weight");
  ws->candidates.n = 0;
  const int generation = next_generation(ws);
  if (r->num_changed > 0) {
#pragma omp parallel for num_threads(e->threads) schedule(static)
    for (int64_t i = 0; i < r->num_changed; ++i) {
      const int32_t v = r->changed_to[i];
      if (r->parent[v] == r->changed_from[i]) __atomic_store_n(&state[v], 2, __ATOMIC_RELAXED);
    }
    int cyclic = 0;
#pragma omp parallel num_threads(e->threads)
    {
#pragma omp for schedule(dynamic, 1024)
      for (int64_t v = 0; v < n; ++v) {
        if (__atomic_load_n(&state[v], __ATOMIC_RELAXED) != 0 || __atomic_load_n(&cyclic, __ATOMIC_RELAXED)) continue;
        int32_t u = (int32_t)v;
        int64_t steps = 0;
        while (__atomic_load_n(&state[u], __ATOMIC_RELAXED) == 0 && r->parent[u] >= 0 && steps <= n) {
          u = r->parent[u];
          ++steps;
        }
        if (steps > n) {
          __atomic_store_n(&cyclic, 1, __ATOMIC_RELAXED);
          continue;
        }
        const char su = __atomic_load_n(&state[u], __ATOMIC_RELAXED);
        const char result = su == 0 ? 1 : su;
        for (u = (int32_t)v; u >= 0 && __atomic_load_n(&state[u], __ATOMIC_RELAXED) == 0; u = r->parent[u])
          __atomic_store_n(&state[u], result, __ATOMIC_RELAXED);
      }
      ivec* local = &ws->tl[2 * omp_get_thread_num()];
      local->n = 0;
#pragma omp for schedule(static) nowait
      for (int64_t v = 0; v < n; ++v) {
        if (state[v] == 2) {
          packed[v] = PACKED_INF;
          stamp[v] = generation;
          ivec_push(local, (int32_t)v);
        }
      }
      gather(&ws->candidates, ws->tl, 0, e->offsets[0]);
    }
    if (cyclic) return fail(err, errlen, ORC_INVALID, "sssp: the input shortest-path tree has a parent cycle");
  }
  r->invalidated = ws->candidates.n;
  if (r->num_insert_heads > 0) {
#pragma omp parallel num_threads(e->threads)
    {
      ivec* local = &ws->tl[2 * omp_get_thread_num()];
      local->n = 0;
#pragma omp for schedule(static) nowait
      for (int64_t i = 0; i < r->num_insert_heads; ++i) {
        const int32_t v = r->insert_heads[i];
        if (v != r->source && claim(stamp, v, generation)) ivec_push(local, v);
      }
      gather(&ws->candidates, ws->tl, 0, e->offsets[0]);
    }
  }
  return ORC_OK;
}

static void omp_seed(omp_engine* e) {
  if (e->empty) return;
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  ws->frontier.n = 0;
  const int pull_generation = next_generation(ws);
  const int64_t count = ws->candidates.n;
  const int32_t* cand = ws->candidates.v;
  u64* packed = ws->packed;
  int* stamp = ws->stamp;
  const packing pk = e->pk;
#pragma omp parallel num_threads(e->threads)
  {
    ivec* local = &ws->tl[2 * omp_get_thread_num()];
    local->n = 0;
#pragma omp for schedule(dynamic, 64) nowait
    for (int64_t i = 0; i < count; ++i) {
      const int32_t v = cand[i];
      const u64 current = load64(&packed[v]);
      u64 best = current;
      for (int64_t q = r->in_row[v]; q < r->in_row[v + 1]; ++q) {
        const int32_t u = r->in_col[q];
        const u64 word = load64(&packed[u]);
        if (word != PACKED_INF) {
          const u64 c = pk_pack(pk, pk_dist(pk, word) + (u64)r->in_w[q], u);
          if (c < best) best = c;
        }
      }
      if (best < current) {
        const u64 old = atomic_min64(&packed[v], best);
        if (pk_dist(pk, best) < pk_dist(pk, old) && claim(stamp, v, pull_generation)) ivec_push(local, v);
      }
    }
    gather(&ws->frontier, ws->tl, 0, e->offsets[0]);
  }
}

static u64 smallest_distance(const ivec* list, const u64* packed, packing pk, int threads) {
  u64 smallest = PACKED_INF;
  const int64_t count = list->n;
#pragma omp parallel for num_threads(threads) reduction(min : smallest) schedule(static)
  for (int64_t i = 0; i < count; ++i) {
    const u64 word = load64(&packed[list->v[i]]);
    if (word != PACKED_INF) {
      const u64 d = pk_dist(pk, word);
      if (d < smallest) smallest = d;
    }
  }
  return smallest;
}

static void omp_loop(omp_engine* e) {
  if (e->empty) return;
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  if (ws->frontier.n == 0) return;
  const u64 delta = (u64)r->delta;
  const packing pk = e->pk;
  u64* packed = ws->packed;
  int* stamp = ws->stamp;
  char* in_far = ws->in_far;
  const u64 smallest = smallest_distance(&ws->frontier, packed, pk, e->threads);
  u64 threshold = (smallest == PACKED_INF ? 0 : smallest) + delta;
  ivec* current = &ws->near_a;
  ivec* next = &ws->near_b;
  current->n = 0;
  ws->far.n = 0;
  {
    const int generation = next_generation(ws);
    const int64_t count = ws->frontier.n;
    const int32_t* frontier = ws->frontier.v;
#pragma omp parallel num_threads(e->threads)
    {
      ivec* ln = &ws->tl[2 * omp_get_thread_num()];
      ivec* lf = &ws->tl[2 * omp_get_thread_num() + 1];
      ln->n = lf->n = 0;
#pragma omp for schedule(static) nowait
      for (int64_t i = 0; i < count; ++i) {
        const int32_t v = frontier[i];
        const u64 word = load64(&packed[v]);
        const u64 d = word == PACKED_INF ? PACKED_INF : pk_dist(pk, word);
        if (d < threshold) {
          if (claim(stamp, v, generation)) ivec_push(ln, v);
        } else if (__atomic_exchange_n(&in_far[v], 1, __ATOMIC_RELAXED) == 0) {
          ivec_push(lf, v);
        }
      }
      gather(current, ws->tl, 0, e->offsets[0]);
      gather(&ws->far, ws->tl, 1, e->offsets[1]);
    }
  }
  for (;;) {
    while (current->n > 0) {
      ++r->iterations;
      r->pushes += current->n;
      const int generation = next_generation(ws);
      const int64_t count = current->n;
      const int32_t* near = current->v;
      next->n = 0;
#pragma omp parallel num_threads(e->threads)
      {
        ivec* ln = &ws->tl[2 * omp_get_thread_num()];
        ivec* lf = &ws->tl[2 * omp_get_thread_num() + 1];
        ln->n = lf->n = 0;
#pragma omp for schedule(dynamic, 64) nowait
        for (int64_t i = 0; i < count; ++i) {
          const int32_t u = near[i];
          const u64 word = load64(&packed[u]);
          if (word == PACKED_INF) continue;
          const u64 du = pk_dist(pk, word);
          for (int64_t q = r->out_row[u]; q < r->out_row[u + 1]; ++q) {
            const int32_t w = r->out_col[q];
            if (w == r->source) continue;
            const u64 nd = du + (u64)r->out_w[q];
            const u64 cand = pk_pack(pk, nd, u);
            if (cand >= load64(&packed[w])) continue;
            const u64 old = atomic_min64(&packed[w], cand);
            if (nd < pk_dist(pk, old)) {
              if (nd < threshold) {
                if (claim(stamp, w, generation)) ivec_push(ln, w);
              } else if (__atomic_exchange_n(&in_far[w], 1, __ATOMIC_RELAXED) == 0) {
                ivec_push(lf, w);
              }
            }
          }
        }
        gather(next, ws->tl, 0, e->offsets[0]);
        gather(&ws->far, ws->tl, 1, e->offsets[1]);
      }
      ivec* t = current;
      current = next;
      next = t;
    }
    if (ws->far.n == 0) break;
    ++r->epochs;
    const u64 sm = smallest_distance(&ws->far, packed, pk, e->threads);
    threshold = (threshold > sm ? threshold : sm) + delta;
    const int generation = next_generation(ws);
    const int64_t count = ws->far.n;
    const int32_t* far = ws->far.v;
    current->n = 0;
    ws->far2.n = 0;
#pragma omp parallel num_threads(e->threads)
    {
      ivec* ln = &ws->tl[2 * omp_get_thread_num()];
      ivec* lk = &ws->tl[2 * omp_get_thread_num() + 1];
      ln->n = lk->n = 0;
#pragma omp for schedule(static) nowait
      for (int64_t i = 0; i < count; ++i) {
        const int32_t v = far[i];
        const u64 word = load64(&packed[v]);
        const u64 d = word == PACKED_INF ? PACKED_INF : pk_dist(pk, word);
        if (d < threshold) {
          in_far[v] = 0;
          if (claim(stamp, v, generation)) ivec_push(ln, v);
        } else {
          ivec_push(lk, v);
        }
      }
      gather(current, ws->tl, 0, e->offsets[0]);
      gather(&ws->far2, ws->tl, 1, e->offsets[1]);
    }
    ivec t = ws->far;
    ws->far = ws->far2;
    ws->far2 = t;
  }
  /* leave the swapped lists in the workspace */
  if (current != &ws->near_a) {
    ivec t = ws->near_a;
    ws->near_a = ws->near_b;
    ws->near_b = t;
  }
}

static void omp_finalize(omp_engine* e) {
  if (e->empty) return;
  orc_run* r = e->run;
  orc_ws* ws = e->ws;
  const int64_t n = r->n;
  const packing pk = e->pk;
  const u64* packed = ws->packed;
  int64_t affected = 0;
  if (pk.bits > 0) {
#pragma omp parallel for num_threads(e->threads) schedule(static) reduction(+ : affected)
    for (int64_t v = 0; v < n; ++v) {
      const u64 word = packed[v];
      const int64_t d = word == PACKED_INF ? ORC_INF : (int64_t)pk_dist(pk, word);
      const int32_t p = word == PACKED_INF ? -1 : pk_parent(pk, word);
      affected += (d != r->dist[v] || p != r->parent[v]) ? 1 : 0;
      r->dist[v] = d;
      r->parent[v] = p;
    }
    r->affected = affected;
    return;
  }
  int32_t* saved = ws->saved_parents;
  char* distance_changed = ws->state;
#pragma omp parallel for num_threads(e->threads) schedule(static)
  for (int64_t v = 0; v < n; ++v) {
    const u64 word = packed[v];
    const int64_t d = word == PACKED_INF ? ORC_INF : (int64_t)word;
    distance_changed[v] = d != r->dist[v] ? 1 : 0;
    saved[v] = r->parent[v];
    r->dist[v] = d;
    r->parent[v] = (word == PACKED_INF || v == r->source) ? -1 : INT32_MAX;
  }
#pragma omp parallel for num_threads(e->threads) schedule(dynamic, 256)
  for (int64_t u = 0; u < n; ++u) {
    const u64 du = packed[u];
    if (du == PACKED_INF) continue;
    for (int64_t q = r->out_row[u]; q < r->out_row[u + 1]; ++q) {
      const int32_t w = r->out_col[q];
      if (w != r->source && du + (u64)r->out_w[q] == packed[w]) atomic_min_id(&r->parent[w], (int32_t)u);
    }
  }
#pragma omp parallel for num_threads(e->threads) schedule(static) reduction(+ : affected)
  for (int64_t v = 0; v < n; ++v) affected += (distance_changed[v] != 0 || saved[v] != r->parent[v]) ? 1 : 0;
  r->affected = affected;
}

int orc_omp_compute(orc_run* run, orc_ws* ws_in, char* err, int errlen) {
  orc_ws* ws = ws_in ? ws_in : orc_ws_new();
  omp_engine* e = (omp_engine*)calloc(1, sizeof(omp_engine));
  int rc = omp_init(e, run, ws, err, errlen);
  if (rc == ORC_OK && !e->empty && (run->source < 0 || run->source >= run->n))
    rc = fail(err, errlen, ORC_INVALID, "sssp::compute: source %d is out of range [0, %lld)", run->source, (long long)run->n);
  if (rc == ORC_OK && !e->empty) {
    /* bind_static assigns INF / -1 before the engine runs (sssp.cpp:812-813) */
    for (int64_t v = 0; v < run->n; ++v) {
      run->dist[v] = ORC_INF;
      run->parent[v] = -1;
    }
    double t0 = now_s();
    /* reset + seed_static (openmp.cpp:197-218) */
    const packing pk = e->pk;
#pragma omp parallel for num_threads(e->threads) schedule(static)
    for (int64_t v = 0; v < run->n; ++v) ws->packed[v] = v == run->source ? pk_pack(pk, 0, -1) : PACKED_INF;
    ws->frontier.n = 0;
    ivec_push(&ws->frontier, run->source);
    double t1 = now_s();
    omp_loop(e);
    double t2 = now_s();
    omp_finalize(e);
    run->t_hooks[0] = 0;
    run->t_hooks[1] = t1 - t0;
    run->t_hooks[2] = t2 - t1;
    run->t_hooks[3] = now_s() - t2;
  }
  free(e);
  if (!ws_in) orc_ws_free(ws);
  return rc;
}

int orc_omp_update(orc_run* run, orc_ws* ws_in, char* err, int errlen) {
  orc_ws* ws = ws_in ? ws_in : orc_ws_new();
  omp_engine* e = (omp_engine*)calloc(1, sizeof(omp_engine));
  int rc = omp_init(e, run, ws, err, errlen);
  if (rc == ORC_OK) {
    double t0 = now_s();
    rc = omp_identify_affected(e, err, errlen);
    double t1 = now_s();
    if (rc == ORC_OK) {
      omp_seed(e);
      double t2 = now_s();
      omp_loop(e);
      double t3 = now_s();
      omp_finalize(e);
      run->t_hooks[0] = t1 - t0;
      run->t_hooks[1] = t2 - t1;
      run->t_hooks[2] = t3 - t2;
      run->t_hooks[3] = now_s() - t3;
    }
  }
  free(e);
  if (!ws_in) orc_ws_free(ws);
  return rc;
}

/* ------------------------------------------------------------------------------------------------
 * Dijkstra with lowest-id ties (cpp/src/testing/dijkstra.cpp:28-71), binary heap.
 * ----------------------------------------------------------------------------------------------*/
typedef struct heap_node {
  int64_t d;
  int32_t v;
} heap_node;

static int node_less(heap_node a, heap_node b) { return a.d < b.d || (a.d == b.d && a.v < b.v); }

void orc_dijkstra(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                  int32_t source, int64_t* dist, int32_t* parent) {
  for (int64_t v = 0; v < n; ++v) {
    dist[v] = ORC_INF;
    parent[v] = -1;
  }
  if (source < 0 || source >= n) return;
  int64_t cap = 1024, size = 0;
  heap_node* h = (heap_node*)malloc((size_t)cap * sizeof(heap_node));
  dist[source] = 0;
  h[size++] = (heap_node){0, source};
  while (size > 0) {
    heap_node top = h[0];
    h[0] = h[--size];
    for (int64_t i = 0;;) {
      int64_t l = 2 * i + 1, r = l + 1, s = i;
      if (l < size && node_less(h[l], h[s])) s = l;
      if (r < size && node_less(h[r], h[s])) s = r;
      if (s == i) break;
      heap_node t = h[i];
      h[i] = h[s];
      h[s] = t;
      i = s;
    }
    if (top.d != dist[top.v]) continue;
    const int32_t u = top.v;
    for (int64_t e = row[u]; e < row[u + 1]; ++e) {
      const int32_t v = col[e];
      const int64_t cand = top.d + (int64_t)w[e];
      if (cand < dist[v]) {
        dist[v] = cand;
        parent[v] = u;
        if (size == cap) {
          cap *= 2;
          h = (heap_node*)realloc(h, (size_t)cap * sizeof(heap_node));
        }
        int64_t i = size++;
        h[i] = (heap_node){cand, v};
        while (i > 0 && node_less(h[i], h[(i - 1) / 2])) {
          heap_node t = h[i];
          h[i] = h[(i - 1) / 2];
          h[(i - 1) / 2] = t;
          i = (i - 1) / 2;
        }
      } else if (cand == dist[v] && v != source && u < parent[v]) {
        parent[v] = u;
      }
    }
  }
  free(h);
}

void orc_check_tree(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                    int32_t source, const int64_t* dist, const int32_t* parent,
                    int32_t require_canonical, int64_t counts[4]) {
  int64_t* rd = (int64_t*)malloc((size_t)(n ? n : 1) * sizeof(int64_t));
  int32_t* rp = (int32_t*)malloc((size_t)(n ? n : 1) * sizeof(int32_t));
  orc_dijkstra(n, row, col, w, source, rd, rp);
  const int64_t m = n > 0 ? row[n] : 0;
  int64_t* in_start = (int64_t*)calloc((size_t)n + 2, sizeof(int64_t));
  for (int64_t e = 0; e < m; ++e) in_start[col[e] + 1]++;
  for (int64_t v = 0; v < n; ++v) in_start[v + 1] += in_start[v];
  int32_t* in_from = (int32_t*)malloc((size_t)(m ? m : 1) * sizeof(int32_t));
  int64_t* in_weight = (int64_t*)malloc((size_t)(m ? m : 1) * sizeof(int64_t));
  int64_t* cursor = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
  memcpy(cursor, in_start, (size_t)(n + 1) * sizeof(int64_t));
  for (int64_t u = 0; u < n; ++u)
    for (int64_t e = row[u]; e < row[u + 1]; ++e) {
      const int32_t v = col[e];
      in_from[cursor[v]] = (int32_t)u;
      in_weight[cursor[v]++] = w[e];
    }
  counts[0] = counts[1] = counts[2] = counts[3] = 0;
  for (int64_t v = 0; v < n; ++v) {
    const int64_t dv = dist[v];
    const int reachable = dv < ORC_INF / 2;
    if (reachable != (rd[v] < ORC_INF / 2) || (reachable && dv != rd[v])) ++counts[0];
    if (require_canonical && parent[v] != rp[v]) ++counts[3];
    if (v == source) {
      if (parent[v] != -1 || dv != 0) ++counts[1];
      continue;
    }
    if (!reachable) {
      if (parent[v] != -1) ++counts[1];
      continue;
    }
    const int32_t p = parent[v];
    int consistent = 0;
    int32_t lowest = -1;
    for (int64_t e = in_start[v]; e < in_start[v + 1]; ++e) {
      const int32_t u = in_from[e];
      if (!(dist[u] < ORC_INF / 2)) continue;
      const int64_t we = in_weight[e];
      if (we <= 0 || dist[u] + we != dv) continue;
      if (u == p) consistent = 1;
      if (lowest < 0 || u < lowest) lowest = u;
    }
    if (!consistent) {
      ++counts[1];
    } else if (p != lowest) {
      ++counts[2];
    }
  }
  free(rd);
  free(rp);
  free(in_start);
  free(in_from);
  free(in_weight);
  free(cursor);
}

void orc_canonicalize(int64_t n, const int64_t* row, const int32_t* col, const int32_t* w,
                      int32_t source, const int64_t* dist, int32_t* parent) {
  for (int64_t u = 0; u < n; ++u) {
    const int64_t du = dist[u];
    if (du >= ORC_INF / 2) continue;
    for (int64_t e = row[u]; e < row[u + 1]; ++e) {
      const int32_t v = col[e];
      if (v != source && u < parent[v] && du + (int64_t)w[e] == dist[v]) parent[v] = (int32_t)u;
    }
  }
}
