/*
 * oracle.c - CPU restatement of gpujoin's index-join probe path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library (libgj.so) never calls
 * it and never falls back to it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository).  Where the reference algorithm lives in an
 * absent third-party dependency (learnedsystems/RadixSpline, unpinned HEAD,
 * lib/RadixSpline.cmake:9-16; fast-interconnects @84c181a,
 * lib/fast-interconnects.cmake:21-32) the published algorithm is restated and
 * the reference call sites are cited.
 *
 * Row ids are uint32 with ORC_NOT_FOUND = 0xFFFFFFFF (numeric_limits<value_t>::max(),
 * src/indexes.cuh:99).  Every lookup checks key equality (the reference's
 * binary_search / radix_spline skip it, src/indexes.cuh:357-361, 205-215;
 * hits agree, misses become ORC_NOT_FOUND).
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   lower_bound / branchy search  brute force on tiny arrays, std-style KAT of src/bf_search.cpp:168-186
 *   B-tree (page layout)          closed-form page arithmetic (leaf 329/247, inner 246), == lower_bound on every key
 *   Harmonia                      reference KATs test/test_harmonia.cu:18-55,115-123 and test/main.cu:144-164
 *   RadixSpline                   invariant |estimate - position| <= max_error for every build key; knot set
 *                                 itself is "parity unpinned" (no reference test holds the third-party output)
 *   hash table                    == lower_bound on hits/misses; murmur3 finaliser constants from the source
 *   partitioner                   brute-force histogram; layout inside the third-party kernel "parity unpinned"
 *   join-aggregate                brute-force sums; closed forms on constructed inputs
 *   Zipf                          bit-equal to the reference's zipf.hpp compiled into oracle/_ref (same mt19937 seed)
 *   MLP gradients                 central finite differences in double
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_NOT_FOUND 0xFFFFFFFFu

static inline uint64_t kat(const void* a, uint64_t i, int kb) {
  return kb == 8 ? ((const uint64_t*)a)[i] : (uint64_t)((const uint32_t*)a)[i];
}
static inline uint64_t kmax(int kb) { return kb == 8 ? ~0ull : 0xFFFFFFFFull; }

/* ------------------------------------------------------------------------
 * Search primitives
 * ------------------------------------------------------------------------ */

/* lower_bound, src/search.cuh:26-41 */
uint64_t orc_lower_bound(const void* arr, uint64_t size, int kb, uint64_t key) {
  uint64_t lower = 0, count = size;
  while (count > 0) {
    uint64_t step = count / 2, mid = lower + step;
    if (kat(arr, mid, kb) < key) {
      lower = mid + 1;
      count -= step + 1;
    } else {
      count = step;
    }
  }
  return lower;
}

/* branchy_binary_search, src/search.cuh:86-101 (reads arr[0] even for size 0,
 * as the reference does; callers pad). */
uint64_t orc_branchy_binary_search(const void* arr, uint64_t size, int kb, uint64_t x) {
  uint64_t lower = 0, upper = size;
  do {
    uint64_t mid = ((upper - lower) / 2) + lower;
    uint64_t v = kat(arr, mid, kb);
    if (x < v) upper = mid;
    else if (x > v) lower = mid + 1;
    else return mid;
  } while (lower < upper);
  return lower;
}

/* binary_search_index::cooperative_lookup (src/indexes.cuh:357-361) with an
 * equality check.  OpenMP over probes; also the CPU baseline (port). */
void orc_lookup_lower_bound(const void* col, uint64_t n, int kb, const void* probes, uint64_t m, uint32_t* out,
                            int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  if (kb == 8) {
    const uint64_t* c = (const uint64_t*)col;
    const uint64_t* p = (const uint64_t*)probes;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t i = 0; i < m; ++i) {
      uint64_t x = p[i], lower = 0, count = n;
      while (count > 0) {
        uint64_t step = count / 2, mid = lower + step;
        if (c[mid] < x) { lower = mid + 1; count -= step + 1; } else { count = step; }
      }
      out[i] = (lower < n && c[lower] == x) ? (uint32_t)lower : ORC_NOT_FOUND;
    }
  } else {
    const uint32_t* c = (const uint32_t*)col;
    const uint32_t* p = (const uint32_t*)probes;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (uint64_t i = 0; i < m; ++i) {
      uint32_t x = p[i];
      uint64_t lower = 0, count = n;
      while (count > 0) {
        uint64_t step = count / 2, mid = lower + step;
        if (c[mid] < x) { lower = mid + 1; count -= step + 1; } else { count = step; }
      }
      out[i] = (lower < n && c[lower] == x) ? (uint32_t)lower : ORC_NOT_FOUND;
    }
  }
}

/* u64 checksum of row ids (config 1 output), wrapping add. */
uint64_t orc_checksum(const uint32_t* rowids, uint64_t m) {
  uint64_t s = 0;
  for (uint64_t i = 0; i < m; ++i) s += rowids[i];
  return s;
}

/* ------------------------------------------------------------------------
 * Clustered B+-tree with the reference's 4 KiB page layout
 * (src/btree.cuh:20-279, 504-531).  Leaves are chunks of the sorted column of
 * LeafNode::maxEntries keys; inner nodes hold separators = max key of each
 * child and maxEntries + 1 children.
 * ------------------------------------------------------------------------ */
typedef struct orc_btree {
  const void* col;
  uint64_t n;
  int kb, vb;
  uint32_t leaf_max, inner_max;
  uint64_t nodes, cap;
  uint32_t* count;
  uint8_t* final_inner;
  uint64_t* keys;     /* nodes x inner_max, padded with the max key (create_inner, src/btree.cuh:124-127) */
  uint64_t* children; /* nodes x (inner_max + 1): node id, or column position below a final inner node */
  uint64_t root;
  uint32_t depth;
} orc_btree;

/* LeafNode/InnerNode::maxEntries, src/btree.cuh:30, 43, 51, 59 (page 4096 B, header 128 B, pointer 8 B). */
void orc_btree_capacities(int kb, int vb, uint32_t* leaf_max, uint32_t* inner_max) {
  *leaf_max = (uint32_t)((4096 - 128) / (kb + vb)) - 1;
  *inner_max = (uint32_t)((4096 - 128 - 8) / (kb + 8)) - 1;
}

static uint64_t bt_new_inner(orc_btree* t, int final_inner) {
  if (t->nodes == t->cap) {
    uint64_t nc = t->cap ? 2 * t->cap : 64;
    t->count = realloc(t->count, nc * sizeof(uint32_t));
    t->final_inner = realloc(t->final_inner, nc);
    t->keys = realloc(t->keys, nc * t->inner_max * sizeof(uint64_t));
    t->children = realloc(t->children, nc * (t->inner_max + 1) * sizeof(uint64_t));
    t->cap = nc;
  }
  uint64_t id = t->nodes++;
  t->count[id] = 0;
  t->final_inner[id] = (uint8_t)final_inner;
  for (uint32_t i = 0; i < t->inner_max; ++i) t->keys[id * t->inner_max + i] = kmax(t->kb);
  for (uint32_t i = 0; i <= t->inner_max; ++i) t->children[id * (t->inner_max + 1) + i] = 0;
  return id;
}

/* max_key, src/btree.cuh:152-165.  is_leaf: the pointer is a column chunk. */
static uint64_t bt_max_key(const orc_btree* t, uint64_t ref, int is_leaf) {
  if (is_leaf) {
    uint64_t cnt = t->n - ref < t->leaf_max ? t->n - ref : t->leaf_max;
    return kat(t->col, ref + cnt - 1, t->kb);
  }
  uint64_t c = t->children[ref * (t->inner_max + 1) + t->count[ref]];
  return bt_max_key(t, c, t->final_inner[ref]);
}

/* construct_inner_nodes, src/btree.cuh:176-206 (load factor 1.0, clustered). */
static uint64_t bt_construct_inner(orc_btree* t, uint64_t* lower, uint64_t lower_n, int lower_is_leaf,
                                   int inverse_depth) {
  if (lower_n == 1 && inverse_depth > 1) return lower[0];
  uint64_t* cur = malloc(sizeof(uint64_t) * (lower_n + 1));
  uint64_t cur_n = 0;
  uint64_t node = bt_new_inner(t, inverse_depth == 1);
  for (uint64_t i = 0; i + 1 < lower_n; ++i) {
    uint64_t curr = lower[i];
    uint64_t sep = bt_max_key(t, curr, lower_is_leaf);
    int full = t->count[node] >= t->inner_max; /* count/maxEntries > 1.0 never holds */
    if (full) {
      t->children[node * (t->inner_max + 1) + t->count[node]] = curr;
      cur[cur_n++] = node;
      node = bt_new_inner(t, inverse_depth == 1);
    } else {
      t->keys[node * t->inner_max + t->count[node]] = sep;
      t->children[node * (t->inner_max + 1) + t->count[node]] = curr;
      t->count[node] += 1;
    }
  }
  t->children[node * (t->inner_max + 1) + t->count[node]] = lower[lower_n - 1];
  cur[cur_n++] = node;
  t->depth += 1;
  uint64_t r = bt_construct_inner(t, cur, cur_n, 0, inverse_depth + 1);
  free(cur);
  return r;
}

/* btree::construct, src/btree.cuh:227-279 (clustered: leaves are &keys[i] every maxEntries keys). */
orc_btree* orc_btree_build(const void* col, uint64_t n, int kb, int vb) {
  if (n == 0) return NULL;
  orc_btree* t = calloc(1, sizeof(orc_btree));
  t->col = col;
  t->n = n;
  t->kb = kb;
  t->vb = vb;
  orc_btree_capacities(kb, vb, &t->leaf_max, &t->inner_max);
  uint64_t leaves_n = (n + t->leaf_max - 1) / t->leaf_max;
  uint64_t* leaves = malloc(sizeof(uint64_t) * leaves_n);
  for (uint64_t i = 0; i < leaves_n; ++i) leaves[i] = i * t->leaf_max;
  t->root = bt_construct_inner(t, leaves, leaves_n, 1, 1);
  free(leaves);
  return t;
}

void orc_btree_info(const orc_btree* t, uint32_t* leaf_max, uint32_t* inner_max, uint32_t* depth, uint64_t* pages) {
  *leaf_max = t->leaf_max;
  *inner_max = t->inner_max;
  *depth = t->depth;
  *pages = t->nodes;
}

/* Clustered lookup, src/btree.cuh:504-531 (equality-checked). */
uint32_t orc_btree_lookup_one(const orc_btree* t, uint64_t key) {
  uint64_t node = t->root;
  uint64_t pos_ref;
  for (;;) {
    uint64_t pos = orc_branchy_binary_search(&t->keys[node * t->inner_max], t->count[node], 8, key);
    uint64_t child = t->children[node * (t->inner_max + 1) + pos];
    if (t->final_inner[node]) {
      pos_ref = child;
      break;
    }
    node = child;
  }
  uint64_t cnt = t->n - pos_ref < t->leaf_max ? t->n - pos_ref : t->leaf_max;
  const uint8_t* base = (const uint8_t*)t->col + pos_ref * (uint64_t)t->kb;
  uint64_t pos = orc_branchy_binary_search(base, cnt, t->kb, key);
  if (pos < cnt && kat(base, pos, t->kb) == key) return (uint32_t)(pos_ref + pos);
  return ORC_NOT_FOUND;
}

void orc_btree_lookup(const orc_btree* t, const void* probes, uint64_t m, uint32_t* out, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
  for (uint64_t i = 0; i < m; ++i) out[i] = orc_btree_lookup_one(t, kat(probes, i, t->kb));
}

void orc_btree_free(orc_btree* t) {
  if (!t) return;
  free(t->count);
  free(t->final_inner);
  free(t->keys);
  free(t->children);
  free(t);
}

/* ------------------------------------------------------------------------
 * Harmonia (src/harmonia.cuh:48-404, 640-668), generic fanout, sorted-only,
 * leaf keys = the input column (Reuse_Input_Memory).
 * ------------------------------------------------------------------------ */
typedef struct orc_level {
  int is_leaf;
  int64_t key_count, node_count, node_prefix, key_prefix;
} orc_level;

typedef struct orc_harmonia {
  const void* col;
  uint64_t n;
  int kb;
  uint64_t fanout, max_keys;
  int depth;
  orc_level* levels;
  uint64_t inner_size; /* keys in the inner levels */
  int reuse_input;     /* 1: leaf keys are the column (Reuse_Input_Memory); 0: copied into keys, padded */
  uint64_t* keys;      /* inner keys (and leaf keys when !reuse_input), padded with the max key */
  uint64_t node_count;
  uint64_t* children;
} orc_harmonia;

static uint64_t hm_get_key(const orc_harmonia* h, uint64_t idx) {
  /* get_key, src/harmonia.cuh:219-230 */
  if (h->reuse_input && idx >= h->inner_size) return kat(h->col, idx - h->inner_size, h->kb);
  return h->keys[idx];
}

/* get_largest_key, src/harmonia.cuh:245-259 */
static uint64_t hm_largest(const orc_harmonia* h, int level_idx, int64_t node_idx) {
  const orc_level* ld = &h->levels[level_idx];
  int64_t nkc = (node_idx < ld->node_count - 1) ? (int64_t)h->max_keys : (ld->key_count - node_idx * (int64_t)h->max_keys);
  if (level_idx == h->depth - 1) {
    int64_t start = (int64_t)h->max_keys * (ld->node_prefix + node_idx);
    return hm_get_key(h, (uint64_t)(start + nkc - 1));
  }
  int64_t child = (int64_t)h->fanout * node_idx + nkc;
  return hm_largest(h, level_idx + 1, child);
}

/* construct, src/harmonia.cuh:346-404.  reuse_input = 1 is the default
 * configuration (leaf keys alias the column); it returns NULL when n is not a
 * multiple of max_keys (src/harmonia.cuh:350-352).  reuse_input = 0 copies the
 * leaf keys into the key array (populate_leaf_nodes, :301-305).  NULL also
 * when the tree exceeds max_depth. */
orc_harmonia* orc_harmonia_build_mode(const void* col, uint64_t n, int kb, uint32_t fanout, int reuse_input) {
  if (n == 0 || fanout < 3) return NULL;
  uint64_t mk = fanout - 1;
  if (reuse_input && n % mk != 0) return NULL;
  orc_harmonia* h = calloc(1, sizeof(orc_harmonia));
  h->reuse_input = reuse_input;
  h->col = col;
  h->n = n;
  h->kb = kb;
  h->fanout = fanout;
  h->max_keys = mk;
  /* gather_tree_info, src/harmonia.cuh:133-185 (levels collected bottom-up, stored top-down) */
  orc_level tmp[64];
  int cnt = 0;
  int64_t prev_pages = 0, kc = (int64_t)n;
  for (;;) {
    int64_t pc = (kc + (int64_t)mk - 1) / (int64_t)mk;
    if ((int64_t)fanout * pc < prev_pages) pc += 1; /* rightmost page empty but with one child */
    tmp[cnt].key_count = kc;
    tmp[cnt].node_count = pc;
    tmp[cnt].is_leaf = 0;
    cnt++;
    prev_pages = pc;
    kc = ((int64_t)mk * pc) / (int64_t)fanout;
    if (pc == 1 || cnt == 64) break;
  }
  if (cnt > 16) { /* max_depth, src/harmonia.cuh:62, 398 */
    free(h);
    return NULL;
  }
  h->depth = cnt;
  h->levels = calloc((size_t)cnt, sizeof(orc_level));
  for (int i = 0; i < cnt; ++i) h->levels[i] = tmp[cnt - 1 - i];
  h->levels[cnt - 1].is_leaf = 1;
  int64_t tn = 0, tk = 0;
  for (int i = 0; i < cnt; ++i) {
    h->levels[i].node_prefix = tn;
    tn += h->levels[i].node_count;
    h->levels[i].key_prefix = tk;
    tk += h->levels[i].key_count;
  }
  const orc_level* leaf = &h->levels[cnt - 1];
  h->inner_size = mk * (uint64_t)leaf->node_prefix;
  h->node_count = (uint64_t)(leaf->node_prefix + leaf->node_count);
  uint64_t key_slots = reuse_input ? h->inner_size : mk * h->node_count;
  h->keys = malloc(sizeof(uint64_t) * (key_slots ? key_slots : 1));
  for (uint64_t i = 0; i < key_slots; ++i) h->keys[i] = kmax(kb);
  if (!reuse_input)
    for (uint64_t i = 0; i < n; ++i) h->keys[h->inner_size + i] = kat(col, i, kb);
  h->children = calloc(h->node_count, sizeof(uint64_t));
  /* store_structure, src/harmonia.cuh:324-344 */
  uint64_t off = 0, prefix = 1;
  for (int l = 0; l < cnt; ++l) {
    const orc_level* lv = &h->levels[l];
    if (lv->is_leaf) break;
    for (int64_t i = 0; i < lv->node_count; ++i) {
      h->children[off++] = prefix;
      int64_t nkc = (i < lv->node_count - 1) ? (int64_t)mk : (lv->key_count - i * (int64_t)mk);
      prefix += (uint64_t)(nkc + 1);
    }
  }
  /* populate_leaf_nodes, src/harmonia.cuh:297-315 */
  off = (uint64_t)leaf->node_prefix;
  uint64_t vps = 0;
  for (int64_t i = 0; i < leaf->node_count; ++i) {
    h->children[off++] = vps;
    vps += mk;
  }
  /* populate_inner_nodes, src/harmonia.cuh:261-295 (from the level above the leaves up to the root) */
  for (int l = cnt - 2; l >= 0; --l) {
    const orc_level* cur = &h->levels[l];
    const orc_level* low = &h->levels[l + 1];
    uint64_t start = (l > 0) ? mk * (uint64_t)cur->node_prefix : 0;
    int skipped = 1;
    uint64_t next = 0;
    for (int64_t i = 0; i < low->node_count - 1; ++i) {
      if (!skipped && (next % mk) == 0) {
        skipped = 1;
        continue;
      }
      h->keys[start + next] = hm_largest(h, l + 1, i);
      next++;
      skipped = 0;
    }
  }
  return h;
}

orc_harmonia* orc_harmonia_build(const void* col, uint64_t n, int kb, uint32_t fanout) {
  return orc_harmonia_build_mode(col, n, kb, fanout, 1);
}

int orc_harmonia_depth(const orc_harmonia* h) { return h->depth; }
uint64_t orc_harmonia_inner_size(const orc_harmonia* h) { return h->inner_size; }
uint64_t orc_harmonia_node_count(const orc_harmonia* h) { return h->node_count; }
void orc_harmonia_dump(const orc_harmonia* h, uint64_t* keys_out, uint64_t* children_out) {
  if (keys_out) memcpy(keys_out, h->keys, sizeof(uint64_t) * h->inner_size);
  if (children_out) memcpy(children_out, h->children, sizeof(uint64_t) * h->node_count);
}

/* host lookup, src/harmonia.cuh:640-668 (the key past a node's end is never
 * read: the reference reads node_start[max_keys] there, which is out of
 * bounds for the last leaf). */
uint32_t orc_harmonia_lookup_one(const orc_harmonia* h, uint64_t key) {
  uint64_t pos = 0, lb = 0;
  int have_actual = 0;
  uint64_t actual = 0;
  for (int d = 1; d <= h->depth; ++d) {
    uint64_t start = h->max_keys * pos;
    lb = h->max_keys;
    for (uint64_t j = 0; j < h->max_keys; ++j) {
      if (hm_get_key(h, start + j) >= key) {
        lb = j;
        break;
      }
    }
    have_actual = lb < h->max_keys;
    if (have_actual) actual = hm_get_key(h, start + lb);
    pos = h->children[pos] + lb;
  }
  if (pos < h->n && have_actual && actual == key) return (uint32_t)pos;
  return ORC_NOT_FOUND;
}

void orc_harmonia_lookup(const orc_harmonia* h, const void* probes, uint64_t m, uint32_t* out, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
  for (uint64_t i = 0; i < m; ++i) out[i] = orc_harmonia_lookup_one(h, kat(probes, i, h->kb));
}

void orc_harmonia_free(orc_harmonia* h) {
  if (!h) return;
  free(h->levels);
  free(h->keys);
  free(h->children);
  free(h);
}

/* ------------------------------------------------------------------------
 * RadixSpline (learnedsystems/RadixSpline builder, restated from the
 * published method: GreedySplineCorridor of Neumann & Michel, BNCOD'08, plus
 * a radix table over the top radix_bits bits of key - min_key).  Reference
 * call sites: build src/rs.cuh:54-64; estimate src/rs.cuh:118-154; search
 * bound src/indexes.cuh:205-215.
 * ------------------------------------------------------------------------ */
typedef struct orc_rs {
  int kb;
  uint64_t min_key, max_key, n;
  uint32_t radix_bits, shift_bits, max_error;
  uint32_t* radix; /* (max_prefix + 2) entries */
  uint64_t radix_n;
  uint64_t* px;    /* knot keys */
  double* py;      /* knot positions */
  uint64_t pn, pcap;
} orc_rs;

static void rs_push(orc_rs* r, uint64_t x, double y, uint32_t* prev_prefix) {
  if (r->pn == r->pcap) {
    r->pcap = r->pcap ? 2 * r->pcap : 1024;
    r->px = realloc(r->px, r->pcap * sizeof(uint64_t));
    r->py = realloc(r->py, r->pcap * sizeof(double));
  }
  r->px[r->pn] = x;
  r->py[r->pn] = y;
  r->pn++;
  /* radix table: every prefix up to this knot's prefix starts at this knot */
  uint32_t cur = (uint32_t)((x - r->min_key) >> r->shift_bits);
  if (cur != *prev_prefix) {
    for (uint64_t p = (uint64_t)*prev_prefix + 1; p <= cur; ++p) r->radix[p] = (uint32_t)(r->pn - 1);
    *prev_prefix = cur;
  }
}

/* sign of the cross product (dx1,dy1) x (dx2,dy2) with a tolerance of one ulp at 1.0 */
static int rs_orient(double dx1, double dy1, double dx2, double dy2) {
  double e = fma(dy1, dx2, -fma(dy2, dx1, 0.0));
  if (e > 2.220446049250313e-16) return 1;   /* clockwise */
  if (e < -2.220446049250313e-16) return -1; /* counter-clockwise */
  return 0;
}

orc_rs* orc_rs_build(const void* col, uint64_t n, int kb, uint32_t radix_bits, uint32_t max_error) {
  if (n == 0) return NULL;
  orc_rs* r = calloc(1, sizeof(orc_rs));
  r->kb = kb;
  r->n = n;
  r->min_key = kat(col, 0, kb);
  r->max_key = kat(col, n - 1, kb);
  r->radix_bits = radix_bits;
  r->max_error = max_error;
  uint64_t diff = r->max_key - r->min_key;
  unsigned used = diff ? (64u - (unsigned)__builtin_clzll(diff)) : 0u; /* significant bits of diff */
  r->shift_bits = (used < radix_bits) ? 0u : used - radix_bits;
  uint64_t max_prefix = diff >> r->shift_bits;
  r->radix_n = max_prefix + 2;
  r->radix = calloc(r->radix_n, sizeof(uint32_t));
  uint32_t prev_prefix = 0;
  double err = (double)max_error;
  /* corridor state */
  uint64_t ux = 0, lx = 0, qx = 0; /* upper/lower limit and previous CDF point keys */
  double uy = 0, ly = 0, qy = 0;
  uint64_t distinct = 0, prev_key = r->min_key;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t key = kat(col, i, kb);
    double pos = (double)i;
    if (i == 0) {
      rs_push(r, key, pos, &prev_prefix);
      distinct = 1;
      qx = key;
      qy = pos;
      prev_key = key;
      continue;
    }
    if (key == prev_key) continue; /* duplicate: not a new CDF point */
    prev_key = key;
    distinct++;
    double upper_y = pos + err;
    double lower_y = (pos < err) ? 0.0 : pos - err;
    if (distinct == 2) {
      ux = key; uy = upper_y;
      lx = key; ly = lower_y;
      qx = key; qy = pos;
      continue;
    }
    uint64_t bx = r->px[r->pn - 1];
    double by = r->py[r->pn - 1];
    double ux_d = (double)(ux - bx), lx_d = (double)(lx - bx), x_d = (double)(key - bx);
    double uy_d = uy - by, ly_d = ly - by, y_d = pos - by;
    if (rs_orient(ux_d, uy_d, x_d, y_d) != 1 || rs_orient(lx_d, ly_d, x_d, y_d) != -1) {
      /* the new point leaves the corridor: the previous CDF point becomes a knot */
      rs_push(r, qx, qy, &prev_prefix);
      ux = key; uy = upper_y;
      lx = key; ly = lower_y;
    } else {
      if (rs_orient(ux_d, uy_d, x_d, upper_y - by) == 1) { ux = key; uy = upper_y; }
      if (rs_orient(lx_d, ly_d, x_d, lower_y - by) == -1) { lx = key; ly = lower_y; }
    }
    qx = key;
    qy = pos;
  }
  /* the last key is always a knot, at the position of its last occurrence */
  if (r->px[r->pn - 1] != prev_key) rs_push(r, prev_key, (double)(n - 1), &prev_prefix);
  /* finalise the radix table: prefixes past the last knot point one past it */
  for (uint64_t p = (uint64_t)prev_prefix + 1; p < r->radix_n; ++p) r->radix[p] = (uint32_t)r->pn;
  return r;
}

uint64_t orc_rs_num_points(const orc_rs* r) { return r->pn; }
uint64_t orc_rs_radix_size(const orc_rs* r) { return r->radix_n; }
uint32_t orc_rs_shift_bits(const orc_rs* r) { return r->shift_bits; }
void orc_rs_dump(const orc_rs* r, uint64_t* px, double* py, uint32_t* radix) {
  if (px) memcpy(px, r->px, r->pn * sizeof(uint64_t));
  if (py) memcpy(py, r->py, r->pn * sizeof(double));
  if (radix) memcpy(radix, r->radix, r->radix_n * sizeof(uint32_t));
}

/* get_spline_segment + get_estimate, src/rs.cuh:118-154 (interpolation with an fma). */
double orc_rs_estimate(const orc_rs* r, uint64_t key) {
  if (key <= r->min_key) return 0.0;
  if (key >= r->max_key) return (double)(r->n - 1);
  uint64_t prefix = (key - r->min_key) >> r->shift_bits;
  uint64_t b = r->radix[prefix], e = r->radix[prefix + 1];
  uint64_t lower = b, count = e - b;
  while (count > 0) { /* lower_bound over knot x */
    uint64_t step = count / 2, mid = lower + step;
    if (r->px[mid] < key) { lower = mid + 1; count -= step + 1; } else { count = step; }
  }
  uint64_t idx = lower;
  double xd = (double)(r->px[idx] - r->px[idx - 1]);
  double yd = r->py[idx] - r->py[idx - 1];
  double slope = yd / xd;
  double kd = (double)(key - r->px[idx - 1]);
  return fma(kd, slope, r->py[idx - 1]);
}

/* radix_spline_index::cooperative_lookup, src/indexes.cuh:205-215 (equality-checked) */
uint32_t orc_rs_lookup_one(const orc_rs* r, const void* col, uint64_t key) {
  double est = orc_rs_estimate(r, key);
  double err = (double)r->max_error;
  uint64_t begin = (est < err) ? 0 : (uint64_t)(est - err);
  uint64_t end = (est + err + 2 > (double)r->n) ? r->n : (uint64_t)(est + err + 2);
  if (begin > end) begin = end;
  const uint8_t* base = (const uint8_t*)col + begin * (uint64_t)r->kb;
  uint64_t pos = begin + orc_lower_bound(base, end - begin, r->kb, key);
  return (pos < r->n && kat(col, pos, r->kb) == key) ? (uint32_t)pos : ORC_NOT_FOUND;
}

void orc_rs_lookup(const orc_rs* r, const void* col, const void* probes, uint64_t m, uint32_t* out, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(threads)
  for (uint64_t i = 0; i < m; ++i) out[i] = orc_rs_lookup_one(r, col, kat(probes, i, r->kb));
}

/* max |estimate(key_i) - i| over the build keys (the invariant that makes the search bound exact) */
double orc_rs_max_abs_error(const orc_rs* r, const void* col) {
  double worst = 0;
  for (uint64_t i = 0; i < r->n; ++i) {
    double d = fabs(orc_rs_estimate(r, kat(col, i, r->kb)) - (double)i);
    if (d > worst) worst = d;
  }
  return worst;
}

void orc_rs_free(orc_rs* r) {
  if (!r) return;
  free(r->radix);
  free(r->px);
  free(r->py);
  free(r);
}

/* ------------------------------------------------------------------------
 * Linear-probing hash table (src/linear_probing_hashtable.cuh:14-98)
 * ------------------------------------------------------------------------ */

/* murmur3_hash 64-bit finaliser, src/cuda_utils.cuh:85-100 */
uint64_t orc_murmur3_64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}
/* murmur3_hash 32-bit finaliser, src/cuda_utils.cuh:65-78 */
uint32_t orc_murmur3_32(uint32_t k) {
  k ^= k >> 16;
  k *= 0x85ebca6bu;
  k ^= k >> 13;
  k *= 0xc2b2ae35u;
  k ^= k >> 16;
  return k;
}

/* calculate_table_size, src/linear_probing_hashtable.cuh:74-79 (load factor 0.5; 64-bit shift) */
uint64_t orc_hash_capacity(uint64_t n) {
  float hint = (float)n / 0.5f;
  uint64_t bits = (uint64_t)ceil(log2((double)hint));
  return 1ull << bits;
}

/* insert (:37-48) every column key with its position, then lookup (:50-72) every probe.
 * Slots hash the key with the finaliser of its width (murmur3_hash's 32- and 64-bit
 * overloads, src/cuda_utils.cuh:65-100). */
static uint64_t orc_key_hash(uint64_t k, int kb) {
  return kb == 4 ? (uint64_t)orc_murmur3_32((uint32_t)k) : orc_murmur3_64(k);
}
void orc_hash_join(const void* col, uint64_t n, int kb, const void* probes, uint64_t m, uint32_t* out) {
  uint64_t cap = orc_hash_capacity(n);
  uint64_t* tk = malloc(cap * sizeof(uint64_t));
  uint32_t* tv = malloc(cap * sizeof(uint32_t));
  const uint64_t empty = ~0ull;
  for (uint64_t i = 0; i < cap; ++i) tk[i] = empty;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t k = kat(col, i, kb);
    uint64_t s = orc_key_hash(k, kb) & (cap - 1);
    for (;;) {
      if (tk[s] == empty || tk[s] == k) {
        tk[s] = k;
        tv[s] = (uint32_t)i;
        break;
      }
      s = (s + 1) & (cap - 1);
    }
  }
  for (uint64_t j = 0; j < m; ++j) {
    uint64_t k = kat(probes, j, kb);
    uint64_t s = orc_key_hash(k, kb) & (cap - 1);
    uint32_t r = ORC_NOT_FOUND;
    for (;;) {
      if (tk[s] == k && k != empty) { r = tv[s]; break; }
      if (tk[s] == empty) break;
      s = (s + 1) & (cap - 1);
    }
    out[j] = r;
  }
  free(tk);
  free(tv);
}

/* ------------------------------------------------------------------------
 * Radix partition (prefix sum + scatter of run_on_stream,
 * src/index_lookup_partitioning.cu:300-366): partition of k is
 * (k >> ignore_bits) & (fanout - 1) (inferred from the RadixPartitionArgs
 * fields, src/index_lookup_partitioning.cu:135-152).  Stable order.
 * ------------------------------------------------------------------------ */
void orc_partition(const void* keys, uint64_t m, int kb, uint32_t radix_bits, uint32_t ignore_bits,
                   uint64_t* counts, void* out_keys, uint32_t* out_payload) {
  uint64_t fan = 1ull << radix_bits;
  uint64_t* start = calloc(fan, sizeof(uint64_t));
  memset(counts, 0, fan * sizeof(uint64_t));
  for (uint64_t i = 0; i < m; ++i) counts[(kat(keys, i, kb) >> ignore_bits) & (fan - 1)]++;
  uint64_t s = 0;
  for (uint64_t p = 0; p < fan; ++p) {
    start[p] = s;
    s += counts[p];
  }
  for (uint64_t i = 0; i < m; ++i) {
    uint64_t k = kat(keys, i, kb);
    uint64_t p = (k >> ignore_bits) & (fan - 1);
    uint64_t d = start[p]++;
    if (kb == 8) ((uint64_t*)out_keys)[d] = k;
    else ((uint32_t*)out_keys)[d] = (uint32_t)k;
    out_payload[d] = (uint32_t)i;
  }
  free(start);
}

/* ------------------------------------------------------------------------
 * Fact-dimension join-aggregate (TPC-H Q14 shape): the correct predicate of
 * hj_probe_kernel (src/tpch/tpch_14_hj.cu:38-58) / query_14_part_build
 * (src/tpch/cpu_only.cpp:230-275); final value of process_results
 * (src/tpch/tpch_14_common.cu:116-132).
 * ------------------------------------------------------------------------ */
void orc_join_aggregate(const uint32_t* dim_sorted_key, const uint8_t* dim_flag, uint64_t dim_n, const int32_t* fkey,
                        const int32_t* date, const int64_t* price, const int8_t* disc, uint64_t rows, int32_t date_lo,
                        int32_t date_hi, int64_t* num, int64_t* den, uint64_t* matches, int threads) {
  if (threads <= 0) threads = omp_get_max_threads();
  int64_t sn = 0, sd = 0;
  uint64_t mt = 0;
#pragma omp parallel for schedule(static) reduction(+ : sn, sd, mt) num_threads(threads)
  for (uint64_t i = 0; i < rows; ++i) {
    if (date[i] < date_lo || date[i] >= date_hi) continue;
    uint64_t pos = orc_lower_bound(dim_sorted_key, dim_n, 4, (uint64_t)(uint32_t)fkey[i]);
    if (pos >= dim_n || dim_sorted_key[pos] != (uint32_t)fkey[i]) continue;
    int64_t s = price[i] * (100 - (int64_t)disc[i]);
    sd += s;
    if (dim_flag[pos]) sn += s;
    mt += 1;
  }
  *num = sn;
  *den = sd;
  *matches = mt;
}

/* process_results (src/tpch/tpch_14_common.cu:124-127); its int64 products wrap in
 * two's complement for large sums (computed through uint64, so defined here). */
int64_t orc_promo_revenue(int64_t num, int64_t den) {
  const int64_t n = (int64_t)((uint64_t)num * 1000u);
  den /= 1000;
  if (den == 0) return 0;
  return (int64_t)((uint64_t)n * 100u) / den;
}

/* ------------------------------------------------------------------------
 * Zipf by rejection-inversion (zipf_distribution, src/zipf.hpp:13-102) fed by
 * MT19937 (Matsumoto & Nishimura 1998) through the libstdc++ mapping
 * uniform_real_distribution(a, b) = generate_canonical<double, 53>() * (b - a) + a,
 * where generate_canonical combines two 32-bit draws.
 * ------------------------------------------------------------------------ */
typedef struct {
  uint32_t mt[624];
  int idx;
} orc_mt19937;

static void mt_seed(orc_mt19937* g, uint32_t s) {
  g->mt[0] = s;
  for (int i = 1; i < 624; ++i) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
  g->idx = 624;
}
static uint32_t mt_next(orc_mt19937* g) {
  if (g->idx >= 624) {
    for (int i = 0; i < 624; ++i) {
      uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
      g->mt[i] = g->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
    }
    g->idx = 0;
  }
  uint32_t y = g->mt[g->idx++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= y >> 18;
  return y;
}
static double mt_canonical(orc_mt19937* g) {
  double lo = (double)mt_next(g);
  double hi = (double)mt_next(g);
  double r = (lo + hi * 4294967296.0) / 18446744073709551616.0;
  if (r >= 1.0) r = nextafter(1.0, 0.0);
  return r;
}

typedef struct {
  double n, q, h_x1, h_n;
} orc_zipf;

static double z_expxm1bx(double x) {
  return (fabs(x) > 1e-8) ? expm1(x) / x : (1.0 + x / 2.0 * (1.0 + x / 3.0 * (1.0 + x / 4.0)));
}
static double z_log1pxbx(double x) {
  return (fabs(x) > 1e-8) ? log1p(x) / x : 1.0 - x * ((1 / 2.0) - x * ((1 / 3.0) - x * (1 / 4.0)));
}
static double z_H(const orc_zipf* z, double x) {
  double lx = log(x);
  return z_expxm1bx((1.0 - z->q) * lx) * lx;
}
static double z_Hinv(const orc_zipf* z, double x) {
  double t = x * (1.0 - z->q);
  if (t < -1.0) t = -1.0;
  return exp(z_log1pxbx(t) * x);
}
static double z_h(const orc_zipf* z, double x) { return exp(-z->q * log(x)); }
static void z_init(orc_zipf* z, uint64_t n, double q) {
  z->n = (double)n;
  z->q = q;
  z->h_x1 = z_H(z, 1.5) - 1.0;
  z->h_n = z_H(z, z->n + 0.5);
}
/* one trial with u01 in [0, 1); 0 = rejected */
static uint64_t z_trial(const orc_zipf* z, double u01) {
  double u = u01 * (z->h_n - z->h_x1) + z->h_x1;
  double x = z_Hinv(z, u);
  double kd = round(x);
  if (kd < 1.0) kd = 1.0;
  if (kd > z->n) kd = z->n;
  return (u >= z_H(z, kd + 0.5) - z_h(z, kd)) ? (uint64_t)kd : 0;
}

/* count draws from Zipf(q) on {1..n} with an mt19937 seeded by seed */
void orc_zipf_mt19937(uint64_t n, double q, uint32_t seed, uint64_t count, uint64_t* out) {
  orc_zipf z;
  z_init(&z, n, q);
  orc_mt19937 g;
  mt_seed(&g, seed);
  for (uint64_t i = 0; i < count; ++i) {
    uint64_t k = 0;
    while (k == 0) k = z_trial(&z, mt_canonical(&g));
    out[i] = k;
  }
}

/* single trial, for generators that draw their own uniforms */
uint64_t orc_zipf_trial(uint64_t n, double q, double u01) {
  orc_zipf z;
  z_init(&z, n, q);
  return z_trial(&z, u01);
}

/* splitmix64 stream (Steele, Lea & Flood 2014): output i of the generator started at state seed */
uint64_t orc_splitmix64(uint64_t seed, uint64_t i) {
  uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* ------------------------------------------------------------------------
 * Learned-index MLP root (no reference counterpart):
 *   f(x) = sum_j w2_j * ramp(w1_j * x + b1_j) + b2,  ramp(a) = min(max(a, 0), 1),
 *   params [w1(h), b1(h), w2(h), b2]
 *   loss = mean((f(x) - y)^2); gradients accumulated in double.
 * ------------------------------------------------------------------------ */
static double orc_ramp(double a) { return a < 0 ? 0 : (a > 1 ? 1 : a); }

void orc_mlp_forward(const float* p, uint32_t h, const float* x, uint64_t n, float* out) {
  const float *w1 = p, *b1 = p + h, *w2 = p + 2 * h, b2 = p[3 * h];
  for (uint64_t i = 0; i < n; ++i) {
    double f = b2;
    for (uint32_t j = 0; j < h; ++j) f += (double)w2[j] * orc_ramp((double)w1[j] * x[i] + b1[j]);
    out[i] = (float)f;
  }
}

double orc_mlp_loss_grad(const float* p, uint32_t h, const float* x, const float* y, uint64_t n, double* grad) {
  const float *w1 = p, *b1 = p + h, *w2 = p + 2 * h, b2 = p[3 * h];
  memset(grad, 0, sizeof(double) * (3 * h + 1));
  double loss = 0;
  for (uint64_t i = 0; i < n; ++i) {
    double f = b2;
    for (uint32_t j = 0; j < h; ++j) f += (double)w2[j] * orc_ramp((double)w1[j] * x[i] + b1[j]);
    double r = f - y[i];
    loss += r * r;
    double g = 2.0 * r / (double)n;
    grad[3 * h] += g;
    for (uint32_t j = 0; j < h; ++j) {
      double a = (double)w1[j] * x[i] + b1[j];
      grad[2 * h + j] += g * orc_ramp(a);
      if (a > 0 && a < 1) {
        grad[j] += g * w2[j] * x[i];
        grad[h + j] += g * w2[j];
      }
    }
  }
  return loss / (double)n;
}
