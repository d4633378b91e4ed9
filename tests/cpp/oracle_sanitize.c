/* oracle_sanitize.c -- drives the oracle's C restatement (oracle/pgvd_oracle.c)
 * through every entry point on small seeded inputs and edge cases, for a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer (any report
 * aborts).  Checks a few invariants on the way; its purpose is the
 * sanitizers' verdict, not parity (the Python tests hold that).
 *
 *     oracle_sanitize    prints "ok" and exits 0, or aborts in a sanitizer
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pgvd_oracle.h"

static uint64_t g_x = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void) {
  g_x ^= g_x << 13;
  g_x ^= g_x >> 7;
  g_x ^= g_x << 17;
  return g_x;
}
static float frnd(void) { return (float)(rnd() >> 40) / (float)(1u << 24) * 2.0f - 1.0f; }

static int g_bad = 0;
#define EXPECT(c)                                                   \
  do {                                                              \
    if (!(c)) {                                                     \
      fprintf(stderr, "%s:%d: expected %s\n", __FILE__, __LINE__, #c); \
      ++g_bad;                                                      \
    }                                                               \
  } while (0)

/* whole path from float points, then every per-tree reader */
static void whole_path(int64_t n, int dim, int bits, int with_labels, int dupes) {
  float *pts = malloc(sizeof(float) * (size_t)(n ? n : 1) * dim);
  for (int64_t i = 0; i < n * dim; ++i) pts[i] = frnd();
  if (dupes)
    for (int64_t i = n / 2; i < n; ++i)
      for (int k = 0; k < dim; ++k) pts[i * dim + k] = pts[(i - n / 2) * dim + k];
  int64_t *q = malloc(sizeof(int64_t) * (size_t)(n ? n : 1) * dim);
  const int64_t nq = oracle_quantize(pts, n, dim, (int64_t)1 << bits, 0, NULL, NULL, 0, q);
  EXPECT(nq == n || nq == 1);
  uint32_t *labels = NULL;
  if (with_labels) {
    labels = malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) labels[i] = (uint32_t)(rnd() % 5);
  }
  oracle_tree *t = oracle_build(q, nq, dim, bits, labels);
  EXPECT(t != NULL);
  if (t) {
    const int64_t m = oracle_tree_m(t), O = oracle_tree_octree_size(t);
    const int NC = 1 << dim;
    uint64_t *lo = malloc(8 * (size_t)m), *hi = malloc(8 * (size_t)m);
    oracle_tree_keys(t, lo, hi);
    for (int64_t i = 1; i < m; ++i) EXPECT(hi[i - 1] < hi[i] || (hi[i - 1] == hi[i] && lo[i - 1] < lo[i]));
    const int64_t nb = m > 1 ? m - 1 : 1;
    int32_t *left = malloc(4 * (size_t)nb), *lcp_len = malloc(4 * (size_t)nb),
            *parent = malloc(4 * (size_t)nb);
    uint8_t *ll = malloc((size_t)nb), *rl = malloc((size_t)nb);
    uint64_t *clo = malloc(8 * (size_t)nb), *chi = malloc(8 * (size_t)nb);
    if (m > 1) oracle_tree_brt(t, left, ll, rl, clo, chi, lcp_len, parent);
    uint32_t *ls = malloc(4 * (size_t)m), *ps = malloc(4 * (size_t)m);
    oracle_tree_splits(t, ls, ps);
    int32_t *children = malloc(4 * (size_t)(O ? O : 1) * NC), *leaf = malloc(4 * (size_t)(O ? O : 1));
    oracle_tree_octree(t, children, leaf);
    if (with_labels) {
      uint8_t *kf = malloc((size_t)m);
      uint32_t *masks = malloc(4 * (size_t)(O ? O : 1));
      (void)oracle_tree_conflicts(t, kf, masks);
      free(kf);
      free(masks);
    }
    /* point location over the grid, with the octree in its record form */
    int32_t *oct = malloc(4 * (size_t)(O ? O : 1) * (NC + 1));
    for (int64_t i = 0; i < O; ++i) {
      for (int c = 0; c < NC; ++c) oct[i * (NC + 1) + c] = children[i * NC + c];
      oct[i * (NC + 1) + NC] = leaf[i];
    }
    const int64_t nf = 500;
    int64_t *fq = malloc(sizeof(int64_t) * nf * dim);
    for (int64_t i = 0; i < nf * dim; ++i) fq[i] = (int64_t)(rnd() % ((uint64_t)1 << bits));
    int32_t *fp = malloc(4 * nf), *fo = malloc(4 * nf), *fd = malloc(4 * nf), *fl = malloc(4 * nf),
            *fs = malloc(4 * nf);
    if (O > 0) oracle_find_leaf(fq, nf, dim, bits, oct, O, fp, fo, fd, fl, fs);
    /* text form, written and read back */
    if (O > 0) {
      const int64_t len = oracle_octree_format_text(oct, O, dim, NULL);
      char *txt = malloc((size_t)len + 1);
      EXPECT(oracle_octree_format_text(oct, O, dim, txt) == len);
      int32_t *back = malloc(4 * (size_t)O * (NC + 1));
      int64_t O2 = -1;
      EXPECT(oracle_octree_parse_text(txt, len, dim, back, O, &O2) == 0 && O2 == O);
      EXPECT(memcmp(back, oct, 4 * (size_t)O * (NC + 1)) == 0);
      EXPECT(oracle_octree_parse_text(txt, len, dim, back, O ? O - 1 : 0, &O2) != 0 || O == 0);
      EXPECT(oracle_octree_parse_text(txt, len > 3 ? len / 2 : len, dim, back, O, &O2) != 0 ||
             len <= 3);
      free(back);
      free(txt);
    }
    /* cell walks in 2D */
    if (dim == 2 && O > 0) {
      const int64_t ns = 200, cap = 4000;
      float *segs = malloc(sizeof(float) * 4 * ns);
      const float w = (float)((int64_t)1 << bits);
      for (int64_t i = 0; i < 4 * ns; ++i) segs[i] = (frnd() * 0.5f + 0.5f) * (w - 1.0f);
      int32_t *counts = malloc(4 * ns), *status = malloc(4 * ns), *cells = malloc(4 * 5 * cap);
      (void)oracle_cell_walk(segs, ns, bits, oct, O, counts, status, cells, cap);
      int32_t *lab = malloc(4 * ns);
      for (int64_t i = 0; i < ns; ++i) lab[i] = (int32_t)(i / 7);
      int32_t *ci_labels = malloc(4 * 8 * (size_t)O);
      float *ci_segs = malloc(sizeof(float) * 32 * (size_t)O);
      oracle_multi_cells(segs, lab, ns, bits, oct, O, ci_labels, ci_segs, status);
      const int64_t fcap = 20000;
      int32_t *extra = malloc(4 * 2 * fcap);
      float *smp = malloc(sizeof(float) * 2 * fcap), *org = malloc(sizeof(float) * 2 * fcap),
            *lens = malloc(sizeof(float) * fcap);
      int64_t nbox = 0;
      int st = 0;
      (void)oracle_find_multi_cells(ci_labels, ci_segs, O, 1.0f, extra, smp, org, lens, fcap, &nbox,
                                    &st);
      free(extra); free(smp); free(org); free(lens);
      free(ci_labels); free(ci_segs); free(lab);
      free(segs); free(counts); free(status); free(cells);
    }
    free(fq); free(fp); free(fo); free(fd); free(fl); free(fs);
    free(oct);
    free(lo); free(hi); free(left); free(lcp_len); free(parent); free(ll); free(rl); free(clo);
    free(chi); free(ls); free(ps); free(children); free(leaf);
    oracle_tree_free(t);
  }
  free(labels);
  free(q);
  free(pts);
}

/* FitBoxes and its helpers on random and degenerate segment pairs */
static void geometry(void) {
  for (int r = 0; r < 3000; ++r) {
    float a[4], b[4];
    for (int k = 0; k < 4; ++k) {
      a[k] = frnd() * 100.0f;
      b[k] = frnd() * 100.0f;
    }
    if (r % 7 == 0) memcpy(b, a, sizeof a);                       /* identical */
    if (r % 11 == 0) { b[0] = a[2]; b[1] = a[3]; }                 /* shared end */
    if (r % 13 == 0) { a[2] = a[0]; a[3] = a[1]; }                 /* a point */
    float smp[2 * 512], org[2 * 512], lens[512];
    int64_t nb = 0;
    int st = 0;
    if (!oracle_fit_boxes_undefined(a, b, 1.0f))
      (void)oracle_fit_boxes(a, b, 1.0f, smp, org, lens, 256, &nb, &st);
    (void)oracle_multi_intersection(a, a + 2, b, b + 2);
    float out[3 * 8];
    const int32_t origin[2] = {(int32_t)(rnd() % 64), (int32_t)(rnd() % 64)};
    (void)oracle_find_intersections(a, b, origin, 1 + (int)(rnd() % 16), out);
  }
  float v[8] = {0, 0, 1, 1, 0.5f, 0.25f, 1, 0}, o[8];
  const float lo[2] = {0, 0};
  EXPECT(oracle_obj2oct(v, 4, lo, 1.0f, 1 << 12, o) == 0);
}

/* the BRT stage alone, equal neighbours included */
static void brt_stage(void) {
  for (int m = 2; m < 300; m += 17) {
    uint64_t *lo = malloc(8 * (size_t)m), *hi = calloc((size_t)m, 8);
    uint64_t v = 0;
    for (int i = 0; i < m; ++i) {
      v += (rnd() % 3 == 0) ? 0 : 1 + rnd() % 1000;
      lo[i] = v;
    }
    int32_t *left = malloc(4 * (size_t)m), *len = malloc(4 * (size_t)m), *par = malloc(4 * (size_t)m);
    uint8_t *ll = malloc((size_t)m), *rl = malloc((size_t)m);
    uint64_t *clo = malloc(8 * (size_t)m), *chi = malloc(8 * (size_t)m);
    EXPECT(oracle_brt_only(lo, hi, m, 40, left, ll, rl, clo, chi, len, par) == 0);
    free(lo); free(hi); free(left); free(len); free(par); free(ll); free(rl); free(clo); free(chi);
  }
  uint32_t in[100], out[100];
  for (int i = 0; i < 100; ++i) in[i] = (uint32_t)(rnd() % 3);
  oracle_stream_scan(in, out, 100);
  EXPECT(oracle_next_pow2(1) == 8 && oracle_next_pow2(9) == 16);
  EXPECT(oracle_compute_lcp_length(1, 0, 1, 0, 40) == 40);
}

/* malformed text: bad tokens, counts and truncation are errors, not reads
 * past the end */
static void text_errors(void) {
  const char *bad[] = {"", "x", "-1", "2 1 2 3 4 5", "1 a b c d e", "1 1 2 3 4 99999999999",
                       "1 1 2 3 4", "00000000000000000000001 -1 -1 -1 -1 15", "1 +", "1 -",
                       "1 1 2 3 4 5x"};
  int32_t rec[64];
  for (size_t i = 0; i < sizeof bad / sizeof *bad; ++i) {
    int64_t O = -1;
    (void)oracle_octree_parse_text(bad[i], (int64_t)strlen(bad[i]), 2, rec, 4, &O);
  }
}

int main(void) {
  const int dims[] = {2, 3};
  for (int d = 0; d < 2; ++d) {
    const int dim = dims[d];
    const int bit_opts[] = {1, 8, dim == 2 ? 16 : 21, dim == 2 ? 30 : 21, dim == 2 ? 62 : 42};
    for (int b = 0; b < 5; ++b)
      for (int64_t n = 1; n <= 4000; n = n * 3 + 1) {
        whole_path(n, dim, bit_opts[b], b == 2, n > 100 && b == 1);
      }
  }
  whole_path(20000, 2, 16, 1, 1);
  whole_path(20000, 3, 21, 1, 0);
  geometry();
  brt_stage();
  text_errors();
  if (g_bad) {
    fprintf(stderr, "%d expectations failed\n", g_bad);
    return 1;
  }
  printf("ok\n");
  return 0;
}
