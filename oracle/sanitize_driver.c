/*
 * sanitize_driver.c -- runs the oracle's threaded entry points on a generated
 * lattice so that a ThreadSanitizer or AddressSanitizer/UBSan build of the
 * oracle can check them (TEST INFRASTRUCTURE ONLY; built by `make sanitize`,
 * run by tests/test_oracle_sanitizers.py).
 *
 * The lattice: side x side nodes on a jittered 100 m grid, two-way edges
 * between grid neighbours (every tenth one-way in the first build), a few exact
 * ties (unjittered rows). Shortest (closeness + betweenness, 8 threads, Phase 1
 * and Phase 2), OD-weighted betweenness and land-use accessibility (8
 * threads) run over all sources of the first build, segment centrality (which needs every edge's reverse twin) over the
 * second; the output is a checksum of the integer density and count rows.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "cityseer_oracle.h"

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static double rnd(void) {
  rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(rng_state >> 11) / 9007199254740992.0;
}

static og_graph *build(uint32_t side, int with_one_way) {
  const uint32_t n = side * side;
  double *xs = malloc(sizeof(double) * n), *ys = malloc(sizeof(double) * n);
  for (uint32_t r = 0; r < side; ++r)
    for (uint32_t c = 0; c < side; ++c) {
      const int tie_row = (r % 4 == 0);
      xs[r * side + c] = 100.0 * c + (tie_row ? 0.0 : 30.0 * (rnd() - 0.5));
      ys[r * side + c] = 100.0 * r + (tie_row ? 0.0 : 30.0 * (rnd() - 0.5));
    }
  const uint32_t cap = 4 * n;
  uint32_t *src = malloc(4 * cap), *dst = malloc(4 * cap), *key = malloc(4 * cap);
  float *len = malloc(4 * cap), *imp = malloc(4 * cap), *ang = malloc(4 * cap), *sec = malloc(4 * cap);
  int32_t *pk = malloc(4 * cap);
  uint32_t m = 0, pair = 0;
  for (uint32_t r = 0; r < side; ++r)
    for (uint32_t c = 0; c < side; ++c) {
      const uint32_t a = r * side + c;
      const uint32_t nbs[2] = {c + 1 < side ? a + 1 : a, r + 1 < side ? a + side : a};
      for (int k = 0; k < 2; ++k) {
        const uint32_t b = nbs[k];
        if (b == a) continue;
        const float l = (float)hypot(xs[a] - xs[b], ys[a] - ys[b]);
        const int one_way = with_one_way && (pair++ % 10) == 0;
        for (int d = 0; d < (one_way ? 1 : 2); ++d) {
          src[m] = d ? b : a;
          dst[m] = d ? a : b;
          key[m] = 0;
          len[m] = l;
          imp[m] = 1.0f;
          ang[m] = 0.0f;
          sec[m] = NAN;
          pk[m] = -1;
          ++m;
        }
      }
    }
  uint8_t *live = malloc(n);
  float *wt = malloc(4 * n);
  for (uint32_t i = 0; i < n; ++i) {
    live[i] = 1;
    wt[i] = 1.0f;
  }
  og_graph *g = og_graph_new(n, NULL, live, wt, m, src, dst, key, len, imp, ang, sec, pk, 0);
  free(xs); free(ys); free(src); free(dst); free(key); free(len); free(imp); free(ang); free(sec); free(pk);
  free(live); free(wt);
  return g;
}

int main(int argc, char **argv) {
  const uint32_t side = argc > 1 ? (uint32_t)atoi(argv[1]) : 14;
  const uint32_t n = side * side;
  og_graph *g = build(side, 1), *g2 = build(side, 0);
  if (!g || !g2) return 2;
  uint8_t *elig = malloc(n);
  uint32_t *sources = malloc(4 * n);
  float *swt = malloc(4 * n);
  for (uint32_t i = 0; i < n; ++i) {
    elig[i] = 1;
    sources[i] = i;
    swt[i] = 1.0f;
  }

  const uint32_t dists_in[3] = {300, 700, 1500};
  uint32_t dists[3], secs[3];
  float betas[3];
  if (og_pair_thresholds(1.33333f, 0, dists_in, NULL, NULL, 3, 0.01831563888873418f, dists, betas, secs))
    return 3;
  og_params p = {0};
  p.n_thresholds = 3;
  p.distances = dists;
  p.betas = betas;
  p.seconds = secs;
  p.speed_m_s = 1.33333f;
  p.tolerance = 1e-4f;
  p.compute_closeness = 1;
  p.compute_betweenness = 1;
  p.n_sources = n;
  p.sources = sources;
  p.source_wt = swt;
  p.eligible = elig;
  p.track_reach = 1;
  p.betw_scale = 1.0;
  p.angular_scaling_unit = 180.0f;
  p.farness_scaling_offset = 1.0f;
  p.n_threads = 8;

  double checksum = 0.0;
  double *out = calloc((size_t)7 * 3 * n, sizeof(double));
  uint32_t reach[3] = {0, 0, 0};
  for (int pass = 0; pass < 2; ++pass) {
    p.tolerance = pass ? 0.05f : 1e-4f;  /* Phase 1, then Phase 2 */
    if (og_centrality_shortest(g, &p, out, reach)) return 4;
    for (size_t i = 0; i < (size_t)3 * n; ++i) checksum += out[i];  /* density rows: integer sums */
  }
  double *seg = calloc((size_t)4 * 3 * n, sizeof(double));
  p.tolerance = 1e-4f;
  if (og_segment_centrality(g2, &p, seg)) return 5;

  /* OD: each origin sends to the next 5 nodes (mod n) */
  uint32_t *off = malloc(4 * (n + 1)), *odd = malloc(4 * 5 * n);
  float *odw = malloc(4 * 5 * n);
  off[0] = 0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t k = 0; k < 5; ++k) {
      odd[5 * i + k] = (i + 1 + 7 * k) % n;
      odw[5 * i + k] = 1.0f + (float)k;
    }
    off[i + 1] = 5 * (i + 1);
  }
  double *od = calloc((size_t)2 * 3 * n, sizeof(double));
  p.tolerance = 1e-4f;
  if (og_betweenness_od_shortest(g, &p, off, odd, odw, od)) return 6;

  /* accessibility: three entries per node, every fourth sharing a dedupe
   * group with the entry before it, four classes (class -1 every seventh) */
  const uint32_t ne = 3 * n, nc = 4;
  uint32_t *grp = malloc(4 * ne), *anode = malloc(4 * ne), *aent = malloc(4 * ne);
  int32_t *cls = malloc(4 * ne);
  double *adist = malloc(8 * ne);
  uint8_t *live = malloc(n);
  for (uint32_t e = 0; e < ne; ++e) {
    grp[e] = (e % 4 == 3) ? e - 1 : e;
    cls[e] = (e % 7 == 6) ? -1 : (int32_t)(e % nc);
    anode[e] = e / 3;
    aent[e] = e;
    adist[e] = 5.0 + 40.0 * rnd();
  }
  for (uint32_t i = 0; i < n; ++i) live[i] = (i % 5) != 0;
  const float mcws[3] = {1.0f, 1.0f, 1.0f};
  double *cnt = malloc(sizeof(double) * nc * 3 * n), *wtd = malloc(sizeof(double) * nc * 3 * n);
  double *near = malloc(sizeof(double) * nc * n);
  if (og_accessibility(g, ne, grp, cls, nc, ne, anode, aent, adist, 3, dists, betas, mcws, secs[2], 1.33333f, 0,
                       live, 8, cnt, wtd, near))
    return 7;
  double acc_sum = 0.0;
  for (size_t i = 0; i < (size_t)nc * 3 * n; ++i) acc_sum += cnt[i];  /* integer counts */

  printf("density checksum %.1f reach %u %u %u access %.1f\n", checksum, reach[0], reach[1], reach[2], acc_sum);
  free(grp); free(anode); free(aent); free(cls); free(adist); free(live); free(cnt); free(wtd); free(near);
  og_graph_free(g);
  og_graph_free(g2);
  free(elig); free(sources); free(swt); free(out); free(seg); free(off); free(odd); free(odw); free(od);
  return 0;
}
