/*
 * CPU restatement of the engine's synthetic-graph recipe.
 *
 * TEST INFRASTRUCTURE ONLY (see bfs_oracle.c).  The reference generates R-MAT
 * edge lists with networkit's RmatGenerator (scripts/genGraph.py:49-55), a
 * third-party package that is not vendored and not installed here, so its
 * exact instances are unpinned.  The engine instead implements the recipe
 * fixed for its workloads: seeded R-MAT quadrant descent, a seeded vertex
 * permutation, self-loop and duplicate removal, symmetrization, and CSC
 * output in the layout of scripts/generator.py:115-147 (col_ptr over source
 * vertices, sorted destination lists in row_idx).
 *
 * Random numbers: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
 * random numbers: as easy as 1, 2, 3", SC'11).  Pinned by the published
 * Random123 known-answer vectors in tests/test_oracle.py.
 *
 * Stream layout (ctr[3] is the stream tag):
 *   tag 0  R-MAT edge k: ctr = (lo(k), hi(k), j/4, 0), word j%4 drives
 *          quadrant choice j (most significant bit first)
 *   tag 1  permutation key of vertex x: ctr = (x, 0, 0, 1)
 *   tag 2  root candidate k: ctr = (lo(k), hi(k), 0, 2)
 *   tag 3  grid edge g keep/drop: ctr = (lo(g), hi(g), 0, 3)
 *   tag 4  grid long-range edge k: ctr = (k, 0, 0, 4)
 * key = (lo(seed), hi(seed)).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                          uint32_t out[4]) {
  uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static void philox(uint64_t seed, uint32_t a, uint32_t b, uint32_t c,
                   uint32_t tag, uint32_t out[4]) {
  uint32_t ctr[4] = {a, b, c, tag};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  oracle_philox4x32_10(ctr, key, out);
}

/* Generated R-MAT edge list before permutation: n_edges pairs over
   [0, 2^scale).  Thresholds are cumulative quadrant probabilities scaled to
   2^32 (tA = a, tAB = a+b, tABC = a+b+c); a 32-bit draw r picks quadrant
   (0,0) if r < tA, (0,1) if r < tAB, (1,0) if r < tABC, else (1,1). */
void oracle_rmat_edges(int scale, uint64_t n_edges, uint32_t tA, uint32_t tAB,
                       uint32_t tABC, uint64_t seed, uint32_t *src,
                       uint32_t *dst) {
  for (uint64_t k = 0; k < n_edges; k++) {
    uint32_t u = 0, v = 0, r[4] = {0, 0, 0, 0};
    for (int j = 0; j < scale; j++) {
      if ((j & 3) == 0)
        philox(seed, (uint32_t)k, (uint32_t)(k >> 32), (uint32_t)(j >> 2), 0, r);
      uint32_t x = r[j & 3];
      uint32_t bit = 1u << (scale - 1 - j);
      if (x < tA) {
      } else if (x < tAB) {
        v |= bit;
      } else if (x < tABC) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    src[k] = u;
    dst[k] = v;
  }
}

static int bits_for(uint64_t n) {
  int b = 0;
  while (((uint64_t)1 << b) < n) b++;
  return b;
}

static void radix_sort_u64(uint64_t *a, uint64_t n, int key_bits) {
  uint64_t *tmp = malloc(sizeof(uint64_t) * (n ? n : 1));
  uint64_t *src = a, *dst = tmp;
  for (int shift = 0; shift < key_bits; shift += 16) {
    uint64_t *cnt = calloc(65537, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) cnt[((src[i] >> shift) & 0xFFFF) + 1]++;
    for (int i = 0; i < 65536; i++) cnt[i + 1] += cnt[i];
    for (uint64_t i = 0; i < n; i++) dst[cnt[(src[i] >> shift) & 0xFFFF]++] = src[i];
    free(cnt);
    uint64_t *t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, sizeof(uint64_t) * n);
  free(tmp);
}

/* perm[x] = rank of ((key64(x) >> b) << b | x) in ascending order, where
   b = ceil(log2 n) and key64(x) = w0 << 32 | w1 of the tag-1 draw. */
void oracle_permutation(uint32_t n, uint64_t seed, uint32_t *perm) {
  int b = bits_for(n);
  uint64_t *keys = malloc(sizeof(uint64_t) * (n ? n : 1));
  for (uint32_t x = 0; x < n; x++) {
    uint32_t r[4];
    philox(seed, x, 0, 0, 1, r);
    uint64_t k64 = ((uint64_t)r[0] << 32) | r[1];
    keys[x] = b >= 64 ? x : (((k64 >> b) << b) | x);
  }
  radix_sort_u64(keys, n, 64);
  uint64_t m = b >= 64 ? ~0ull : (((uint64_t)1 << b) - 1);
  for (uint32_t i = 0; i < n; i++) perm[keys[i] & m] = i;
  free(keys);
}

/* Builds a symmetric CSC from an undirected edge list: drops self loops,
   emits both directions, sorts by (source, destination), removes duplicates.
   col_ptr has n+1 entries of ptr_bytes (4 or 8); row_idx must hold 2*m
   entries.  Returns nnz. */
uint64_t oracle_build_symmetric_csc(uint32_t n, uint64_t m, const uint32_t *u,
                                    const uint32_t *v, int ptr_bytes,
                                    void *col_ptr, uint32_t *row_idx) {
  uint64_t *keys = malloc(sizeof(uint64_t) * (2 * m + 1));
  uint64_t k = 0;
  for (uint64_t i = 0; i < m; i++) {
    if (u[i] == v[i]) continue;
    keys[k++] = ((uint64_t)u[i] << 32) | v[i];
    keys[k++] = ((uint64_t)v[i] << 32) | u[i];
  }
  radix_sort_u64(keys, k, 64);
  uint64_t nnz = 0;
  for (uint64_t i = 0; i < k; i++)
    if (i == 0 || keys[i] != keys[i - 1]) keys[nnz++] = keys[i];
  uint64_t e = 0;
  for (uint64_t s = 0; s <= n; s++) {
    while (e < nnz && (keys[e] >> 32) < s) e++;
    if (ptr_bytes == 8) ((uint64_t *)col_ptr)[s] = e;
    else ((uint32_t *)col_ptr)[s] = (uint32_t)e;
  }
  for (uint64_t i = 0; i < nnz; i++) row_idx[i] = (uint32_t)keys[i];
  free(keys);
  return nnz;
}

/* Full R-MAT recipe: edges, permutation, symmetric CSC. */
uint64_t oracle_rmat_csc(int scale, uint64_t n_edges, uint32_t tA, uint32_t tAB,
                         uint32_t tABC, uint64_t seed_graph, uint64_t seed_perm,
                         int ptr_bytes, void *col_ptr, uint32_t *row_idx) {
  uint32_t n = 1u << scale;
  uint32_t *src = malloc(sizeof(uint32_t) * n_edges);
  uint32_t *dst = malloc(sizeof(uint32_t) * n_edges);
  uint32_t *perm = malloc(sizeof(uint32_t) * n);
  oracle_rmat_edges(scale, n_edges, tA, tAB, tABC, seed_graph, src, dst);
  oracle_permutation(n, seed_perm, perm);
  for (uint64_t i = 0; i < n_edges; i++) {
    src[i] = perm[src[i]];
    dst[i] = perm[dst[i]];
  }
  uint64_t nnz = oracle_build_symmetric_csc(n, n_edges, src, dst, ptr_bytes,
                                            col_ptr, row_idx);
  free(src); free(dst); free(perm);
  return nnz;
}

/* 2D grid side x side, 4-neighbour edges, vertex id r*side + c.  Grid edge g
   (horizontal edges first, g = r*(side-1) + c; then vertical,
   g = side*(side-1) + r*side + c) is kept when its tag-3 draw w0 >= t_drop.
   n_long extra edges k join (w0*V)>>32 and (w1*V)>>32 of the tag-4 draw.
   Returns nnz of the symmetric CSC. */
uint64_t oracle_grid_csc(uint32_t side, uint32_t t_drop, uint32_t n_long,
                         uint64_t seed, int ptr_bytes, void *col_ptr,
                         uint32_t *row_idx) {
  uint64_t V = (uint64_t)side * side;
  uint64_t G = 2ull * side * (side - 1);
  uint32_t *u = malloc(sizeof(uint32_t) * (G + n_long + 1));
  uint32_t *v = malloc(sizeof(uint32_t) * (G + n_long + 1));
  uint64_t m = 0;
  for (uint64_t g = 0; g < G; g++) {
    uint32_t r[4];
    philox(seed, (uint32_t)g, (uint32_t)(g >> 32), 0, 3, r);
    if (r[0] < t_drop) continue;
    uint64_t a, b;
    uint64_t H = (uint64_t)side * (side - 1);
    if (g < H) {
      uint64_t row = g / (side - 1), col = g % (side - 1);
      a = row * side + col;
      b = a + 1;
    } else {
      uint64_t h = g - H;
      uint64_t row = h / side, col = h % side;
      a = row * side + col;
      b = a + side;
    }
    u[m] = (uint32_t)a;
    v[m] = (uint32_t)b;
    m++;
  }
  for (uint32_t k = 0; k < n_long; k++) {
    uint32_t r[4];
    philox(seed, k, 0, 0, 4, r);
    u[m] = (uint32_t)(((uint64_t)r[0] * V) >> 32);
    v[m] = (uint32_t)(((uint64_t)r[1] * V) >> 32);
    m++;
  }
  uint64_t nnz = oracle_build_symmetric_csc((uint32_t)V, m, u, v, ptr_bytes,
                                            col_ptr, row_idx);
  free(u); free(v);
  return nnz;
}

/* Seeded root choice: candidate k is x = (w0*V)>>32 of the tag-2 draw;
   accepted if eligible[x] and not already taken; stops after n_roots accepts
   or max_tries candidates.  Returns the number of roots found. */
uint32_t oracle_pick_roots(uint32_t V, const uint8_t *eligible,
                           uint32_t n_roots, uint64_t seed, uint64_t max_tries,
                           uint32_t *roots) {
  uint32_t got = 0;
  for (uint64_t k = 0; k < max_tries && got < n_roots; k++) {
    uint32_t r[4];
    philox(seed, (uint32_t)k, (uint32_t)(k >> 32), 0, 2, r);
    uint32_t x = (uint32_t)(((uint64_t)r[0] * V) >> 32);
    if (!eligible[x]) continue;
    int dup = 0;
    for (uint32_t i = 0; i < got && !dup; i++) dup = roots[i] == x;
    if (!dup) roots[got++] = x;
  }
  return got;
}
