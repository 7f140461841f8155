/* Exercises every entry point of the CPU oracle (oracle/bfs_oracle.c,
 * oracle/graphgen_ref.c) on small graphs, for the ASan/UBSan build in
 * tests/test_sanitizers.py.  Cross-checks as it goes: the level-synchronous
 * and FIFO forms agree on every root, every multi-source lane equals the
 * minimum over its sources' single searches, a parent array built from the
 * levels passes the parent check, and uint32 / uint64 col_ptr builds agree.
 * Exit status 0 when every check holds. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int64_t oracle_bfs_levelsync(uint32_t, uint32_t, const void *, int, const uint32_t *, uint32_t,
                             int32_t *, uint64_t *);
int64_t oracle_bfs_queue(uint32_t, uint32_t, const void *, int, const uint32_t *, uint32_t,
                         int32_t *, uint64_t *);
int oracle_msbfs_u8(uint32_t, uint32_t, const void *, int, const uint32_t *, const uint32_t *,
                    const uint16_t *, uint32_t, uint32_t, uint8_t *);
uint64_t oracle_check_parents(uint32_t, uint32_t, const void *, int, const uint32_t *, uint32_t,
                              const int32_t *, const uint32_t *);
void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
uint64_t oracle_rmat_csc(int, uint64_t, uint32_t, uint32_t, uint32_t, uint64_t, uint64_t, int,
                         void *, uint32_t *);
uint64_t oracle_grid_csc(uint32_t, uint32_t, uint32_t, uint64_t, int, void *, uint32_t *);
uint32_t oracle_pick_roots(uint32_t, const uint8_t *, uint32_t, uint64_t, uint64_t, uint32_t *);

static int fails = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      fails++;                                                     \
    }                                                              \
  } while (0)

/* rows < cols: the last vertices are never expanded */
static void check_graph(uint32_t n, const uint32_t *cp, const uint32_t *ri, uint32_t n_rows) {
  int32_t *a = malloc(sizeof(int32_t) * n), *b = malloc(sizeof(int32_t) * n);
  uint32_t *par = malloc(sizeof(uint32_t) * n);
  for (uint32_t r = 0; r < n; r += 1 + n / 40) {
    uint64_t ea = 0, eb = 0;
    CHECK(oracle_bfs_levelsync(n_rows, n, cp, 4, ri, r, a, &ea) >= 0);
    CHECK(oracle_bfs_queue(n_rows, n, cp, 4, ri, r, b, &eb) >= 0);
    CHECK(memcmp(a, b, sizeof(int32_t) * n) == 0 && ea == eb);
    /* a parent for every reached vertex: any expanded in-neighbour one level up */
    for (uint32_t v = 0; v < n; v++) par[v] = a[v] < 0 ? 0xFFFFFFFFu : (v == r ? r : 0xFFFFFFFEu);
    for (uint32_t u = 0; u < n_rows; u++)
      for (uint32_t e = cp[u]; e < cp[u + 1]; e++)
        if (a[u] >= 0 && a[ri[e]] == a[u] + 1 && par[ri[e]] == 0xFFFFFFFEu) par[ri[e]] = u;
    CHECK(oracle_check_parents(n_rows, n, cp, 4, ri, r, a, par) == 0);
  }
  /* 40 searches from 37 sources over 9 lanes, against single searches */
  const uint32_t nb = 9, ns = 37;
  uint32_t src[37];
  uint16_t bits[37];
  for (uint32_t i = 0; i < ns; i++) {
    src[i] = (i * 7919u) % n;
    bits[i] = (uint16_t)(i % nb);
  }
  uint8_t *lv = malloc((size_t)nb * n);
  const int rc = oracle_msbfs_u8(n_rows, n, cp, 4, ri, src, bits, ns, nb, lv);
  CHECK(rc == 0 || rc == -2);
  if (rc == 0) {
    uint8_t *want = malloc((size_t)nb * n);
    memset(want, 0xFF, (size_t)nb * n);
    for (uint32_t i = 0; i < ns; i++) {
      uint64_t ec;
      oracle_bfs_queue(n_rows, n, cp, 4, ri, src[i], a, &ec);
      for (uint32_t v = 0; v < n; v++)
        if (a[v] >= 0 && (uint8_t)a[v] < want[(size_t)bits[i] * n + v])
          want[(size_t)bits[i] * n + v] = (uint8_t)a[v];
    }
    CHECK(memcmp(lv, want, (size_t)nb * n) == 0);
    free(want);
  }
  free(lv);
  free(a);
  free(b);
  free(par);
}

int main(void) {
  const uint32_t ctr[4] = {0, 0, 0, 0}, key[2] = {0, 0};
  uint32_t out[4];
  oracle_philox4x32_10(ctr, key, out);
  CHECK(out[0] == 0x6627e8d5u); /* Random123 known answer, counter and key zero */

  const int scale = 10;
  const uint32_t n = 1u << scale;
  const uint64_t m = 16ull << scale;
  uint32_t *cp = malloc(sizeof(uint32_t) * (n + 1));
  uint64_t *cp64 = malloc(sizeof(uint64_t) * (n + 1));
  uint32_t *ri = malloc(sizeof(uint32_t) * 2 * m), *ri64 = malloc(sizeof(uint32_t) * 2 * m);
  const uint32_t tA = 2448131358u, tAB = 3264175144u, tABC = 4080218931u; /* .57/.19/.19 */
  const uint64_t nnz = oracle_rmat_csc(scale, m, tA, tAB, tABC, 11, 12, 4, cp, ri);
  const uint64_t nnz64 = oracle_rmat_csc(scale, m, tA, tAB, tABC, 11, 12, 8, cp64, ri64);
  CHECK(nnz == nnz64 && nnz <= 2 * m && cp[n] == nnz);
  for (uint32_t i = 0; i <= n; i++) CHECK(cp[i] == cp64[i]);
  CHECK(memcmp(ri, ri64, sizeof(uint32_t) * nnz) == 0);
  check_graph(n, cp, ri, n);
  check_graph(n, cp, ri, n - 5);

  uint8_t *elig = malloc(n);
  for (uint32_t v = 0; v < n; v++) elig[v] = cp[v + 1] > cp[v];
  uint32_t roots[64];
  const uint32_t got = oracle_pick_roots(n, elig, 64, 3001, 64ull * n + 1024, roots);
  CHECK(got == 64);
  for (uint32_t i = 0; i < got; i++) CHECK(elig[roots[i]]);

  const uint32_t side = 40, V = side * side, G = 2 * side * (side - 1), nl = 5;
  uint32_t *gcp = malloc(sizeof(uint32_t) * (V + 1)), *gri = malloc(sizeof(uint32_t) * 2 * (G + nl));
  const uint64_t gnnz = oracle_grid_csc(side, 214748364u, nl, 77, 4, gcp, gri);
  CHECK(gnnz <= 2ull * (G + nl) && gcp[V] == gnnz);
  check_graph(V, gcp, gri, V);

  free(cp);
  free(cp64);
  free(ri);
  free(ri64);
  free(elig);
  free(gcp);
  free(gri);
  if (fails) return 1;
  printf("oracle ok\n");
  return 0;
}
