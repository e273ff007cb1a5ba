/*
 * cityseer_oracle.c -- CPU restatement of cityseer's localised network-centrality
 * path. TEST INFRASTRUCTURE ONLY: parity checker for the HIP kernels and the CPU
 * baseline for bench.py. Never linked into the product library.
 *
 * Structure follows the reference deliberately, so that its cost profile is the
 * reference's (one binary heap per source, node_bound-sized state allocated per
 * source, small multi-predecessor lists, f64 compare-and-swap accumulators,
 * dynamic scheduling of sources over a thread pool).
 *
 * Citations are rust/src/<file>:<line> in the cityseer repository.
 */
#define _GNU_SOURCE
#include "cityseer_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* centrality.rs:24-28 */
#define ANGULAR_ROUTE_TIE_BREAK_FACTOR 1e-6f
#define TIE_EPSILON 1e-4f
/* common.rs:10-12 */
#define MIN_THRESH_WT 0.01831563888873418f
/* thresholds per call (the reference takes any number; tests use up to 20) */
#define OG_MAX_T 32

/* ------------------------------------------------------------------------- */
/* graph: petgraph StableGraph adjacency emulation                           */
/* ------------------------------------------------------------------------- */

struct og_graph {
  uint32_t n, m;
  double *z;
  uint8_t *live;
  float *weight;
  uint32_t *esrc, *edst, *ekey;
  float *elen, *eimp, *eang, *esec;
  int32_t *epk;
  int is_dual;
  /* petgraph keeps per-node singly linked edge lists, newest edge first
   * (Graph::add_edge links the new edge at the head of both lists). */
  int64_t *out_first, *in_first, *out_next, *in_next;
  /* endpoint slots per dual node (centrality.rs:533-566): up to 2 keys. */
  int32_t *slot_key; /* [2N], -1 = empty */
  int slots_ok;
};

static void *xmalloc(size_t n) { return malloc(n ? n : 1); }

og_graph *og_graph_new(uint32_t n_nodes, const double *z, const uint8_t *live,
                       const float *weight, uint32_t n_edges, const uint32_t *src,
                       const uint32_t *dst, const uint32_t *edge_key,
                       const float *length, const float *imp_factor,
                       const float *angle_sum, const float *seconds,
                       const int32_t *primal_key, int is_dual) {
  og_graph *g = (og_graph *)calloc(1, sizeof(og_graph));
  g->n = n_nodes;
  g->m = n_edges;
  g->z = xmalloc(sizeof(double) * n_nodes);
  g->live = xmalloc(n_nodes);
  g->weight = xmalloc(sizeof(float) * n_nodes);
  for (uint32_t i = 0; i < n_nodes; i++) {
    g->z[i] = z ? z[i] : NAN;
    g->live[i] = live ? live[i] : 1;
    g->weight[i] = weight ? weight[i] : 1.0f;
  }
  g->esrc = xmalloc(sizeof(uint32_t) * n_edges);
  g->edst = xmalloc(sizeof(uint32_t) * n_edges);
  g->ekey = xmalloc(sizeof(uint32_t) * n_edges);
  g->elen = xmalloc(sizeof(float) * n_edges);
  g->eimp = xmalloc(sizeof(float) * n_edges);
  g->eang = xmalloc(sizeof(float) * n_edges);
  g->esec = xmalloc(sizeof(float) * n_edges);
  g->epk = xmalloc(sizeof(int32_t) * n_edges);
  for (uint32_t e = 0; e < n_edges; e++) {
    g->esrc[e] = src[e];
    g->edst[e] = dst[e];
    g->ekey[e] = edge_key ? edge_key[e] : 0;
    g->elen[e] = length[e];
    g->eimp[e] = imp_factor ? imp_factor[e] : 1.0f;
    g->eang[e] = angle_sum ? angle_sum[e] : 0.0f;
    g->esec[e] = seconds ? seconds[e] : NAN;
    g->epk[e] = primal_key ? primal_key[e] : -1;
  }
  g->is_dual = is_dual;
  g->out_first = xmalloc(sizeof(int64_t) * n_nodes);
  g->in_first = xmalloc(sizeof(int64_t) * n_nodes);
  g->out_next = xmalloc(sizeof(int64_t) * n_edges);
  g->in_next = xmalloc(sizeof(int64_t) * n_edges);
  for (uint32_t i = 0; i < n_nodes; i++) g->out_first[i] = g->in_first[i] = -1;
  for (uint32_t e = 0; e < n_edges; e++) {
    uint32_t a = src[e], b = dst[e];
    g->out_next[e] = g->out_first[a];
    g->in_next[e] = g->in_first[b];
    g->out_first[a] = e;
    g->in_first[b] = e;
  }
  /* dual_node_endpoint_slots: edges in index order, source then target. */
  g->slot_key = xmalloc(sizeof(int32_t) * 2 * (size_t)n_nodes);
  for (size_t i = 0; i < 2 * (size_t)n_nodes; i++) g->slot_key[i] = -1;
  g->slots_ok = 1;
  for (uint32_t e = 0; e < n_edges && g->slots_ok; e++) {
    int32_t k = g->epk[e];
    if (k < 0) {
      g->slots_ok = 0;
      break;
    }
    uint32_t ends[2] = {src[e], dst[e]};
    for (int t = 0; t < 2; t++) {
      int32_t *s = &g->slot_key[2 * (size_t)ends[t]];
      if (s[0] == k || s[1] == k) continue;
      if (s[0] < 0)
        s[0] = k;
      else if (s[1] < 0)
        s[1] = k;
      else {
        g->slots_ok = 0;
        break;
      }
    }
  }
  return g;
}

void og_graph_free(og_graph *g) {
  if (!g) return;
  free(g->z);
  free(g->live);
  free(g->weight);
  free(g->esrc);
  free(g->edst);
  free(g->ekey);
  free(g->elen);
  free(g->eimp);
  free(g->eang);
  free(g->esec);
  free(g->epk);
  free(g->out_first);
  free(g->in_first);
  free(g->out_next);
  free(g->in_next);
  free(g->slot_key);
  free(g);
}

/* centrality.rs:970-986 */
static float slope_penalty(const og_graph *g, uint32_t from, uint32_t to, float length_2d) {
  if (length_2d <= 0.0f) return 1.0f;
  double zf = g->z[from], zt = g->z[to];
  if (isnan(zf) || isnan(zt)) return 1.0f;
  float slope = (float)(zt - zf) / length_2d;
  const float FLAT_FACTOR = 0.839457f;
  float slope_factor = expf(-3.5f * fabsf(slope + 0.05f));
  return FLAT_FACTOR / slope_factor;
}

/* centrality.rs:989-1007 */
static float edge_travel_seconds(const og_graph *g, uint32_t from, uint32_t to, uint32_t e,
                                 float speed, int use_impedance) {
  if (!isnan(g->esec[e])) return g->esec[e];
  float sp = slope_penalty(g, from, to, g->elen[e]);
  float imp = use_impedance ? g->eimp[e] : 1.0f;
  float a = g->elen[e] * imp;
  a = a * sp;
  return a / speed;
}

/* graph.rs:1278-1292 (edges_connecting(start, end).find(edge_idx)) */
static int64_t edge_connecting(const og_graph *g, uint32_t start, uint32_t end, uint32_t key) {
  for (int64_t e = g->out_first[start]; e >= 0; e = g->out_next[e])
    if (g->edst[e] == end && g->ekey[e] == key) return e;
  return -1;
}

/* ------------------------------------------------------------------------- */
/* thresholds: common.rs:98-270                                              */
/* ------------------------------------------------------------------------- */

static int betas_from_distances(const uint32_t *d, uint32_t n, float mtw, float *out) {
  if (n == 0) return OG_EINVAL;
  for (uint32_t i = 1; i < n; i++)
    if (d[i] <= d[i - 1]) return OG_EINVAL;
  for (uint32_t i = 0; i < n; i++) {
    if (d[i] == 0) return OG_EINVAL;
    float beta = -logf(mtw) / (float)d[i];
    out[i] = roundf(beta * 100000.0f) / 100000.0f;
  }
  return OG_OK;
}

static int distances_from_betas(const float *b, uint32_t n, float mtw, uint32_t *out) {
  if (n == 0) return OG_EINVAL;
  for (uint32_t i = 1; i < n; i++)
    if (b[i] >= b[i - 1]) return OG_EINVAL;
  for (uint32_t i = 0; i < n; i++) {
    if (b[i] <= 0.0f) return OG_EINVAL;
    float d = roundf(logf(mtw) / -b[i]);
    if (d <= 0.0f) return OG_EINVAL;
    out[i] = (uint32_t)d;
  }
  return OG_OK;
}

static int seconds_from_distances(const uint32_t *d, uint32_t n, float speed, uint32_t *out) {
  if (n == 0 || speed <= 0.0f) return OG_EINVAL;
  for (uint32_t i = 1; i < n; i++)
    if (d[i] <= d[i - 1]) return OG_EINVAL;
  for (uint32_t i = 0; i < n; i++) {
    if (d[i] == 0) return OG_EINVAL;
    float t = roundf((float)d[i] / speed);
    if (t <= 0.0f) return OG_EINVAL;
    out[i] = (uint32_t)t;
  }
  return OG_OK;
}

static int distances_from_seconds(const uint32_t *s, uint32_t n, float speed, uint32_t *out) {
  if (n == 0 || speed <= 0.0f) return OG_EINVAL;
  for (uint32_t i = 1; i < n; i++)
    if (s[i] <= s[i - 1]) return OG_EINVAL;
  for (uint32_t i = 0; i < n; i++) {
    if (s[i] == 0) return OG_EINVAL;
    float d = roundf((float)s[i] * speed);
    if (d <= 0.0f) return OG_EINVAL;
    out[i] = (uint32_t)d;
  }
  return OG_OK;
}

int og_pair_thresholds(float speed, int mode, const uint32_t *distances, const float *betas,
                       const float *minutes, uint32_t n, float mtw, uint32_t *od, float *ob,
                       uint32_t *os) {
  int rc;
  if (mode == 0) {
    for (uint32_t i = 0; i < n; i++) od[i] = distances[i];
    if ((rc = betas_from_distances(od, n, mtw, ob))) return rc;
    return seconds_from_distances(od, n, speed, os);
  } else if (mode == 1) {
    for (uint32_t i = 0; i < n; i++) ob[i] = betas[i];
    if ((rc = distances_from_betas(ob, n, mtw, od))) return rc;
    return seconds_from_distances(od, n, speed, os);
  } else if (mode == 2) {
    for (uint32_t i = 0; i < n; i++) {
      float s = roundf(minutes[i] * 60.0f);
      os[i] = (uint32_t)s;
    }
    if ((rc = distances_from_seconds(os, n, speed, od))) return rc;
    return betas_from_distances(od, n, mtw, ob);
  }
  return OG_EINVAL;
}

/* ------------------------------------------------------------------------- */
/* priority queues -- centrality.rs:358-386                                  */
/*                                                                           */
/* Mode 0 (keyed): min-heap on (metric, aux, idx); equal metrics pop in a    */
/* deterministic (seconds, index) order that a parallel search reproduces.   */
/* Mode 1 (rust): the algorithm of Rust's std BinaryHeap (push = sift_up;    */
/* pop = swap-remove, sift_down_to_bottom, sift_up) over NodeDistance, whose */
/* Ord compares the metric only (reversed total_cmp). It reproduces the      */
/* reference's pop order among equal metrics exactly, given the same pushes. */
/* On tie-free inputs both modes pop in the same order.                      */
/* ------------------------------------------------------------------------- */

static int g_heap_mode = 0;

void og_set_heap_mode(int rust_order) { g_heap_mode = rust_order ? 1 : 0; }

/* Planted mistakes (og_set_mutation), 0 = none. The tests switch one on to show
 * that a pinning check fails on a plausible error, then switch it off again.
 * Nothing else sets them. */
static int g_mutation = 0;

void og_set_mutation(int m) { g_mutation = m; }

typedef struct {
  float k;
  float aux;
  uint32_t idx;
} hentry;

typedef struct {
  hentry *a;
  size_t n, cap;
} heap;

static int hless(hentry x, hentry y) {
  if (x.k != y.k) return x.k < y.k;
  if (x.aux != y.aux) return x.aux < y.aux;
  return x.idx < y.idx;
}

/* NodeDistance ordering: a <= b in Rust's Ord  <=>  a.metric >= b.metric */
static int rust_le(hentry a, hentry b) { return a.k >= b.k; }

static size_t rust_sift_up(heap *h, size_t start, size_t pos) {
  hentry el = h->a[pos];
  while (pos > start) {
    size_t parent = (pos - 1) / 2;
    if (rust_le(el, h->a[parent])) break;
    h->a[pos] = h->a[parent];
    pos = parent;
  }
  h->a[pos] = el;
  return pos;
}

static void hgrow(heap *h) {
  if (h->n == h->cap) {
    h->cap = h->cap ? h->cap * 2 : 64;
    h->a = realloc(h->a, h->cap * sizeof(hentry));
  }
}

static void hpush(heap *h, float k, float aux, uint32_t idx) {
  hgrow(h);
  if (g_heap_mode) {
    h->a[h->n] = (hentry){k, aux, idx};
    h->n++;
    rust_sift_up(h, 0, h->n - 1);
    return;
  }
  size_t i = h->n++;
  hentry e = {k, aux, idx};
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (!hless(e, h->a[p])) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = e;
}

static int hpop(heap *h, hentry *out) {
  if (h->n == 0) return 0;
  if (g_heap_mode) {
    hentry item = h->a[--h->n];
    if (h->n > 0) {
      hentry t = h->a[0];
      h->a[0] = item;
      item = t;
      /* sift_down_to_bottom(0) */
      size_t end = h->n, pos = 0, child = 1;
      hentry el = h->a[0];
      while (end >= 2 && child <= end - 2) {
        child += rust_le(h->a[child], h->a[child + 1]) ? 1 : 0;
        h->a[pos] = h->a[child];
        pos = child;
        child = 2 * pos + 1;
      }
      if (child == end - 1) {
        h->a[pos] = h->a[child];
        pos = child;
      }
      h->a[pos] = el;
      rust_sift_up(h, 0, pos);
    }
    *out = item;
    return 1;
  }
  *out = h->a[0];
  hentry last = h->a[--h->n];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, s = i;
    hentry best = last;
    if (l < h->n && hless(h->a[l], best)) {
      s = l;
      best = h->a[l];
    }
    if (r < h->n && hless(h->a[r], best)) {
      s = r;
      best = h->a[r];
    }
    if (s == i) break;
    h->a[i] = h->a[s];
    i = s;
  }
  if (h->n) h->a[i] = last;
  return 1;
}

/* ------------------------------------------------------------------------- */
/* Brandes traversal state: centrality.rs:394-424                            */
/* ------------------------------------------------------------------------- */

typedef struct {
  uint32_t *v;
  uint32_t n, cap;
} plist; /* SmallVec<[usize; 2]> */

static void plist_push(plist *p, uint32_t x) {
  if (p->n == p->cap) {
    p->cap = p->cap ? p->cap * 2 : 2;
    p->v = realloc(p->v, p->cap * sizeof(uint32_t));
  }
  p->v[p->n++] = x;
}
static int plist_has(const plist *p, uint32_t x) {
  for (uint32_t i = 0; i < p->n; i++)
    if (p->v[i] == x) return 1;
  return 0;
}

typedef struct {
  uint32_t n_states;
  uint8_t *visited;
  plist *preds;
  double *sigma;
  float *route;
  float *secs;
  uint32_t *order;
  uint32_t n_order;
  float *best_route; /* [N] */
  float *best_secs;  /* [N] */
} traversal;

static void trav_init(traversal *t, uint32_t n_nodes, uint32_t n_states) {
  t->n_states = n_states;
  t->visited = calloc(n_states, 1);
  t->preds = calloc(n_states, sizeof(plist));
  t->sigma = calloc(n_states, sizeof(double));
  t->route = xmalloc(sizeof(float) * n_states);
  t->secs = xmalloc(sizeof(float) * n_states);
  t->order = xmalloc(sizeof(uint32_t) * n_states);
  t->n_order = 0;
  t->best_route = xmalloc(sizeof(float) * n_nodes);
  t->best_secs = xmalloc(sizeof(float) * n_nodes);
  for (uint32_t i = 0; i < n_states; i++) t->route[i] = t->secs[i] = INFINITY;
  for (uint32_t i = 0; i < n_nodes; i++) t->best_route[i] = t->best_secs[i] = INFINITY;
}

static void trav_free(traversal *t) {
  for (uint32_t i = 0; i < t->n_states; i++) free(t->preds[i].v);
  free(t->visited);
  free(t->preds);
  free(t->sigma);
  free(t->route);
  free(t->secs);
  free(t->order);
  free(t->best_route);
  free(t->best_secs);
}

/* dijkstra_brandes_shortest: centrality.rs:1344-1494 */
static void brandes_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                             float tolerance, traversal *t) {
  const uint32_t N = g->n;
  trav_init(t, N, N);
  const float maxsec = (float)max_seconds;
  t->sigma[src] = 1.0;
  t->route[src] = 0.0f;
  t->secs[src] = 0.0f;
  t->best_route[src] = 0.0f;
  t->best_secs[src] = 0.0f;
  heap h = {0};
  hpush(&h, 0.0f, 0.0f, src);
  hentry top;
  while (hpop(&h, &top)) {
    uint32_t u = top.idx;
    if (t->visited[u]) continue;
    t->visited[u] = 1;
    t->order[t->n_order++] = u;
    for (int64_t e = g->in_first[u]; e >= 0; e = g->in_next[e]) {
      uint32_t nb = g->esrc[e];
      if (nb == u) continue;
      float es = edge_travel_seconds(g, nb, u, (uint32_t)e, speed, 1);
      float cs = t->secs[u] + es;
      if (cs > maxsec) continue;
      if (t->visited[nb]) continue;
      float cr = cs * speed;
      int improved = cs < t->secs[nb];
      int tied = cs <= t->secs[nb] * (1.0f + TIE_EPSILON);
      if (improved) {
        if (cs < t->secs[nb] * (1.0f - TIE_EPSILON)) {
          t->preds[nb].n = 0;
          t->sigma[nb] = t->sigma[u];
        } else {
          t->sigma[nb] += t->sigma[u];
        }
        t->route[nb] = cr;
        t->secs[nb] = cs;
        plist_push(&t->preds[nb], u);
        hpush(&h, cr, cs, nb);
      } else if (tied && !plist_has(&t->preds[nb], u)) {
        if (cs < t->secs[nb]) t->secs[nb] = cs;
        plist_push(&t->preds[nb], u);
        t->sigma[nb] += t->sigma[u];
      }
      t->best_route[nb] = t->route[nb];
      t->best_secs[nb] = t->secs[nb];
    }
  }
  free(h.a);
  if (tolerance > TIE_EPSILON) {
    /* Phase 2: centrality.rs:1447-1483 */
    size_t *pos = xmalloc(sizeof(size_t) * N);
    for (uint32_t i = 0; i < N; i++) pos[i] = (size_t)-1;
    for (uint32_t i = 0; i < t->n_order; i++) pos[t->order[i]] = i;
    for (uint32_t i = 0; i < t->n_order; i++) {
      t->preds[t->order[i]].n = 0;
      t->sigma[t->order[i]] = 0.0;
    }
    t->sigma[src] = 1.0;
    for (uint32_t i = 0; i < t->n_order; i++) {
      uint32_t u = t->order[i];
      for (int64_t e = g->in_first[u]; e >= 0; e = g->in_next[e]) {
        uint32_t v = g->esrc[e];
        if (v == u) continue;
        if (pos[v] == (size_t)-1 || pos[v] <= i) continue;
        float es = edge_travel_seconds(g, v, u, (uint32_t)e, speed, 1);
        float ps = t->secs[u] + es;
        if (ps <= t->secs[v] * (1.0f + tolerance) && !plist_has(&t->preds[v], u)) {
          plist_push(&t->preds[v], u);
          t->sigma[v] += t->sigma[u];
        }
      }
    }
    free(pos);
  }
}

/* dijkstra_brandes_angular: centrality.rs:577-775. Requires slots_ok. */
static int slot_pos(const og_graph *g, uint32_t node, int32_t key) {
  const int32_t *s = &g->slot_key[2 * (size_t)node];
  if (s[0] == key) return 0;
  if (s[1] == key) return 1;
  return -1;
}

static void brandes_angular(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                            float tolerance, traversal *t) {
  const uint32_t N = g->n;
  trav_init(t, N, 2 * N);
  const float maxsec = (float)max_seconds;
  heap h = {0};
  t->best_route[src] = 0.0f;
  t->best_secs[src] = 0.0f;
  for (uint32_t slot = 0; slot < 2; slot++) {
    uint32_t s = 2 * src + slot;
    t->sigma[s] = 1.0;
    t->route[s] = 0.0f;
    t->secs[s] = 0.0f;
    hpush(&h, 0.0f, 0.0f, s);
  }
  hentry top;
  while (hpop(&h, &top)) {
    uint32_t us = top.idx;
    if (t->visited[us]) continue;
    t->visited[us] = 1;
    t->order[t->n_order++] = us;
    uint32_t u = us / 2, entry = us % 2;
    for (int64_t e = g->out_first[u]; e >= 0; e = g->out_next[e]) {
      uint32_t v = g->edst[e];
      int32_t key = g->epk[e];
      int cur_slot = slot_pos(g, u, key);
      if (cur_slot != (int)(1 - entry)) continue;
      int nxt_slot = slot_pos(g, v, key);
      uint32_t vs = 2 * v + (uint32_t)nxt_slot;
      float es = edge_travel_seconds(g, u, v, (uint32_t)e, speed, 0);
      float cs = t->secs[us] + es;
      if (cs > maxsec) continue;
      if (t->visited[vs]) continue;
      float cr = t->route[us] + g->eang[e];
      cr = cr + ANGULAR_ROUTE_TIE_BREAK_FACTOR * g->elen[e];
      float cur = t->route[vs];
      int improved = cr < cur;
      int tied = cr <= cur * (1.0f + TIE_EPSILON);
      if (improved) {
        if (cr < cur * (1.0f - TIE_EPSILON)) {
          t->preds[vs].n = 0;
          t->sigma[vs] = t->sigma[us];
        } else {
          t->sigma[vs] += t->sigma[us];
        }
        t->route[vs] = cr;
        t->secs[vs] = cs;
        plist_push(&t->preds[vs], us);
        hpush(&h, cr, cs, vs);
      } else if (tied && !plist_has(&t->preds[vs], us)) {
        if (cs < t->secs[vs]) t->secs[vs] = cs;
        plist_push(&t->preds[vs], us);
        t->sigma[vs] += t->sigma[us];
      }
      float best = t->best_route[v];
      if (cr < best * (1.0f - TIE_EPSILON)) {
        t->best_route[v] = cr;
        t->best_secs[v] = cs;
      } else if (cr <= best * (1.0f + TIE_EPSILON)) {
        t->best_secs[v] = fminf(t->best_secs[v], cs);
      }
    }
  }
  free(h.a);
  if (tolerance > TIE_EPSILON) {
    /* Phase 2: centrality.rs:710-764 */
    size_t *pos = xmalloc(sizeof(size_t) * 2 * N);
    for (uint32_t i = 0; i < 2 * N; i++) pos[i] = (size_t)-1;
    for (uint32_t i = 0; i < t->n_order; i++) pos[t->order[i]] = i;
    for (uint32_t i = 0; i < t->n_order; i++) {
      t->preds[t->order[i]].n = 0;
      t->sigma[t->order[i]] = 0.0;
    }
    t->sigma[2 * src] = 1.0;
    t->sigma[2 * src + 1] = 1.0;
    for (uint32_t i = 0; i < t->n_order; i++) {
      uint32_t us = t->order[i];
      uint32_t u = us / 2, entry = us % 2;
      for (int64_t e = g->out_first[u]; e >= 0; e = g->out_next[e]) {
        uint32_t v = g->edst[e];
        int32_t key = g->epk[e];
        if (slot_pos(g, u, key) != (int)(1 - entry)) continue;
        uint32_t vs = 2 * v + (uint32_t)slot_pos(g, v, key);
        if (pos[vs] == (size_t)-1 || pos[vs] <= i) continue;
        float cr = t->route[us] + g->eang[e];
        cr = cr + ANGULAR_ROUTE_TIE_BREAK_FACTOR * g->elen[e];
        if (cr <= t->route[vs] * (1.0f + tolerance) && !plist_has(&t->preds[vs], us)) {
          plist_push(&t->preds[vs], us);
          t->sigma[vs] += t->sigma[us];
        }
      }
    }
    free(pos);
  }
}

/* circuit_ranks_from_traversal: centrality.rs:470-526. Distinct undirected
 * edges are keyed (min, max, edge_idx); counting each key once is equivalent
 * to the reference's HashSet de-duplication. */
typedef struct {
  uint32_t a, b, k;
} ekey3;

static int ekey3_cmp(const void *x, const void *y) {
  const ekey3 *p = x, *q = y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  if (p->k != q->k) return p->k < q->k ? -1 : 1;
  return 0;
}

static void circuit_ranks(const og_graph *g, const traversal *t, const uint32_t *reached,
                          uint32_t n_reached, const og_params *p, float *out) {
  const uint32_t T = p->n_thresholds;
  size_t node_counts[OG_MAX_T] = {0}, edge_counts[OG_MAX_T] = {0};
  for (uint32_t r = 0; r < n_reached; r++) {
    float c = t->best_route[reached[r]];
    for (uint32_t i = 0; i < T; i++)
      if (c <= (float)p->distances[i]) node_counts[i]++;
  }
  /* collect (key, cost) for every outgoing edge between reached nodes */
  size_t cap = 64, n = 0;
  ekey3 *keys = xmalloc(sizeof(ekey3) * cap);
  float *costs = xmalloc(sizeof(float) * cap);
  for (uint32_t r = 0; r < n_reached; r++) {
    uint32_t u = reached[r];
    for (int64_t e = g->out_first[u]; e >= 0; e = g->out_next[e]) {
      uint32_t v = g->edst[e];
      if (v == u) continue;
      if (!isfinite(t->best_route[v])) continue;
      if (n == cap) {
        cap *= 2;
        keys = realloc(keys, sizeof(ekey3) * cap);
        costs = realloc(costs, sizeof(float) * cap);
      }
      keys[n].a = u < v ? u : v;
      keys[n].b = u < v ? v : u;
      keys[n].k = g->ekey[e];
      costs[n] = fmaxf(t->best_route[u], t->best_route[v]);
      n++;
    }
  }
  /* the cost of a key is identical for every edge sharing it (max of its two
   * endpoint costs), so counting distinct keys per threshold is exact. */
  typedef struct {
    ekey3 k;
    float c;
  } kc;
  kc *arr = xmalloc(sizeof(kc) * (n ? n : 1));
  for (size_t i = 0; i < n; i++) {
    arr[i].k = keys[i];
    arr[i].c = costs[i];
  }
  qsort(arr, n, sizeof(kc), ekey3_cmp); /* kc starts with ekey3 */
  for (size_t i = 0; i < n; i++) {
    if (i > 0 && ekey3_cmp(&arr[i].k, &arr[i - 1].k) == 0) continue;
    for (uint32_t j = 0; j < T; j++)
      if (arr[i].c <= (float)p->distances[j]) edge_counts[j]++;
  }
  for (uint32_t i = 0; i < T; i++) {
    if (node_counts[i] == 0)
      out[i] = 0.0f;
    else {
      long long r = (long long)edge_counts[i] - (long long)node_counts[i] + 1;
      out[i] = (float)(r > 0 ? r : 0);
    }
  }
  free(arr);
  free(keys);
  free(costs);
}

/* sorted_brandes_state_indices: centrality.rs:777-791 (stable sort by route desc) */
typedef struct {
  float route;
  uint32_t pos;
  uint32_t state;
} sentry;

static int sentry_cmp(const void *x, const void *y) {
  const sentry *a = x, *b = y;
  if (a->route != b->route) return a->route > b->route ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);
}

static uint32_t sorted_states(const traversal *t, uint32_t *out) {
  sentry *s = xmalloc(sizeof(sentry) * (t->n_order ? t->n_order : 1));
  uint32_t n = 0;
  for (uint32_t i = 0; i < t->n_order; i++) {
    uint32_t st = t->order[i];
    if (t->sigma[st] > 0.0) {
      s[n].route = t->route[st];
      s[n].pos = i;
      s[n].state = st;
      n++;
    }
  }
  qsort(s, n, sizeof(sentry), sentry_cmp);
  for (uint32_t i = 0; i < n; i++) out[i] = s[i].state;
  free(s);
  return n;
}

/* f64 accumulator with relaxed CAS add (atomic_float::AtomicF64::fetch_add) */
static void atomic_add_f64(double *addr, double v) {
  _Atomic uint64_t *a = (_Atomic uint64_t *)addr;
  uint64_t old = atomic_load_explicit(a, memory_order_relaxed);
  for (;;) {
    double cur;
    memcpy(&cur, &old, 8);
    double nv = cur + v;
    uint64_t nb;
    memcpy(&nb, &nv, 8);
    if (atomic_compare_exchange_weak_explicit(a, &old, nb, memory_order_relaxed,
                                              memory_order_relaxed))
      return;
  }
}

/* brandes_backprop_with_beta: centrality.rs:823-873. include_mode 0: route <=
 * threshold (shortest), 1: secs <= threshold (angular). state_div = 1 for
 * shortest, 2 for angular (state -> node). */
static void backprop(const traversal *t, const uint32_t *sorted, uint32_t n_sorted, uint32_t src,
                     const double *seed, const double *seed_b, int include_mode, float thr,
                     uint32_t state_div, double *delta, double *delta_b, double *betw,
                     double *betw_b, double wt) {
  memset(delta, 0, sizeof(double) * t->n_states);
  memset(delta_b, 0, sizeof(double) * t->n_states);
  for (uint32_t k = 0; k < n_sorted; k++) {
    uint32_t w = sorted[k];
    int inc = include_mode == 0 ? (t->route[w] <= thr) : (t->secs[w] <= thr);
    if (!inc) continue;
    double sw = t->sigma[w];
    if (sw == 0.0) continue;
    double dep = seed[w] + delta[w];
    double dep_b = seed_b[w] + delta_b[w];
    if (dep == 0.0 && dep_b == 0.0) continue;
    for (uint32_t j = 0; j < t->preds[w].n; j++) {
      uint32_t pv = t->preds[w].v[j];
      double sv = t->sigma[pv];
      if (sv == 0.0) continue;
      double f = sv / sw;
      delta[pv] += f * dep;
      delta_b[pv] += f * dep_b;
    }
    uint32_t node = w / state_div;
    if (node == src) continue;
    double credit = dep - seed[w];
    double credit_b = dep_b - seed_b[w];
    if (credit > 0.0 || credit_b > 0.0) {
      double c = credit > 0.0 ? credit : 0.0;
      double cb = credit_b > 0.0 ? credit_b : 0.0;
      if (c > 0.0) atomic_add_f64(&betw[node], c * wt);
      if (betw_b && cb > 0.0) atomic_add_f64(&betw_b[node], cb * wt);
    }
  }
}

/* ------------------------------------------------------------------------- */
/* thread pool over sources (rayon par_iter stand-in)                        */
/* ------------------------------------------------------------------------- */

typedef struct {
  const og_graph *g;
  const og_params *p;
  double *out;
  _Atomic uint32_t *reach;
  _Atomic uint32_t next;
  int kind; /* 0 shortest, 1 simplest, 2 segment, 3 OD-weighted betweenness */
  _Atomic int err;
  /* kind 3: the OD row of origin p->sources[i] is od_dest/od_w[od_off[i] .. od_off[i + 1]) */
  const uint32_t *od_off;
  const uint32_t *od_dest;
  const float *od_w;
} job;

static uint32_t max_u32(const uint32_t *a, uint32_t n) {
  uint32_t m = 0;
  for (uint32_t i = 0; i < n; i++)
    if (a[i] > m) m = a[i];
  return m;
}

#define OUT(M, I, NODE) (out[((size_t)(M) * T + (I)) * N + (NODE)])

/* centrality_shortest per-source body: centrality.rs:1703-1835 */
static void shortest_source(job *jb, uint32_t s_i) {
  const og_graph *g = jb->g;
  const og_params *p = jb->p;
  double *out = jb->out;
  const uint32_t N = g->n, T = p->n_thresholds;
  uint32_t src = p->sources[s_i];
  if (!p->eligible[src]) return;
  float wt = p->source_wt[s_i];
  if (isnan(wt)) return;
  uint32_t maxs = max_u32(p->seconds, T);
  traversal t;
  brandes_shortest(g, src, maxs, p->speed_m_s, p->tolerance, &t);
  /* reached_node_indices: centrality.rs:460-466 */
  uint32_t *reached = xmalloc(sizeof(uint32_t) * N);
  uint32_t nr = 0;
  for (uint32_t i = 0; i < N; i++)
    if (isfinite(t.best_route[i])) reached[nr++] = i;
  float cycles_wt = wt / g->weight[src];
  if (p->compute_closeness) {
    float score[OG_MAX_T];
    circuit_ranks(g, &t, reached, nr, p, score);
    for (uint32_t r = 0; r < nr; r++) {
      uint32_t to = reached[r];
      if (to == src) continue;
      if (!isfinite(t.best_secs[to])) continue;
      float rc = t.best_route[to];
      if (p->track_reach)
        for (uint32_t i = 0; i < T; i++)
          if (rc <= (float)p->distances[i]) atomic_fetch_add(&jb->reach[i], 1);
      for (uint32_t i = 0; i < T; i++) {
        if (rc <= (float)p->distances[i]) {
          atomic_add_f64(&OUT(0, i, to), (double)wt);
          atomic_add_f64(&OUT(1, i, to), (double)(rc * wt));
          atomic_add_f64(&OUT(2, i, to), (double)(score[i] * cycles_wt));
          atomic_add_f64(&OUT(3, i, to), (double)((1.0f / rc) * wt));
          float nb = -p->betas[i];
          atomic_add_f64(&OUT(4, i, to), (double)(expf(nb * rc) * wt));
        }
      }
    }
  }
  if (p->compute_betweenness) {
    uint32_t *sorted = xmalloc(sizeof(uint32_t) * (t.n_order ? t.n_order : 1));
    uint32_t ns = sorted_states(&t, sorted);
    double *seed = xmalloc(sizeof(double) * N), *seed_b = xmalloc(sizeof(double) * N);
    double *delta = xmalloc(sizeof(double) * N), *delta_b = xmalloc(sizeof(double) * N);
    for (uint32_t i = 0; i < T; i++) {
      float thr = (float)p->distances[i];
      double beta = (double)p->betas[i];
      memset(seed, 0, sizeof(double) * N);
      memset(seed_b, 0, sizeof(double) * N);
      for (uint32_t r = 0; r < nr; r++) {
        uint32_t to = reached[r];
        if (to == src) continue;
        if (t.best_route[to] > thr) continue;
        double pc = p->eligible[to] ? 0.5 : 1.0;
        double pb = pc * exp(-beta * (double)t.best_route[to]);
        if (g_mutation == OG_MUT_BETA_SOURCE_ROUTE) pb = pc * exp(-beta * (double)t.best_route[src]);
        if (g_mutation == OG_MUT_BETA_FULL_SEED) pb = exp(-beta * (double)t.best_route[to]);
        seed[to] += pc;
        seed_b[to] += pb;
      }
      backprop(&t, sorted, ns, src, seed, seed_b, 0, thr, 1, delta, delta_b, &OUT(5, i, 0),
               &OUT(6, i, 0), (double)wt);
    }
    free(sorted);
    free(seed);
    free(seed_b);
    free(delta);
    free(delta_b);
  }
  free(reached);
  trav_free(&t);
}

/* betweenness_od_shortest per-origin body: centrality.rs:2475-2535. Brandes
 * search as for centrality_shortest; per threshold the seeds are the origin's
 * OD weights at destinations within the threshold (w and w * exp(-beta *
 * route)); credits are not scaled by a source weight. Out rows: betweenness,
 * betweenness_beta. */
static void od_source(job *jb, uint32_t s_i) {
  const og_graph *g = jb->g;
  const og_params *p = jb->p;
  double *out = jb->out;
  const uint32_t N = g->n, T = p->n_thresholds;
  uint32_t src = p->sources[s_i];
  if (!g->live[src]) return;
  uint32_t maxs = max_u32(p->seconds, T);
  traversal t;
  brandes_shortest(g, src, maxs, p->speed_m_s, p->tolerance, &t);
  uint32_t *sorted = xmalloc(sizeof(uint32_t) * (t.n_order ? t.n_order : 1));
  uint32_t ns = sorted_states(&t, sorted);
  double *seed = xmalloc(sizeof(double) * N), *seed_b = xmalloc(sizeof(double) * N);
  double *delta = xmalloc(sizeof(double) * N), *delta_b = xmalloc(sizeof(double) * N);
  for (uint32_t i = 0; i < T; i++) {
    float thr = (float)p->distances[i];
    double beta = (double)p->betas[i];
    memset(seed, 0, sizeof(double) * N);
    memset(seed_b, 0, sizeof(double) * N);
    for (uint32_t k = jb->od_off[s_i]; k < jb->od_off[s_i + 1]; k++) {
      uint32_t dest = jb->od_dest[k];
      if (t.best_route[dest] > thr) continue;
      double w = (double)jb->od_w[k];
      seed[dest] += w;
      seed_b[dest] += w * exp(-beta * (double)t.best_route[dest]);
    }
    backprop(&t, sorted, ns, src, seed, seed_b, 0, thr, 1, delta, delta_b, &OUT(0, i, 0), &OUT(1, i, 0), 1.0);
  }
  free(sorted);
  free(seed);
  free(seed_b);
  free(delta);
  free(delta_b);
  trav_free(&t);
}

/* centrality_simplest per-source body: centrality.rs:1967-2097 */
static void simplest_source(job *jb, uint32_t s_i) {
  const og_graph *g = jb->g;
  const og_params *p = jb->p;
  double *out = jb->out;
  const uint32_t N = g->n, T = p->n_thresholds;
  uint32_t src = p->sources[s_i];
  if (!p->eligible[src]) return;
  float wt = p->source_wt[s_i];
  if (isnan(wt)) return;
  uint32_t maxs = max_u32(p->seconds, T);
  traversal t;
  brandes_angular(g, src, maxs, p->speed_m_s, p->tolerance, &t);
  uint32_t *reached = xmalloc(sizeof(uint32_t) * N);
  uint32_t nr = 0;
  for (uint32_t i = 0; i < N; i++)
    if (isfinite(t.best_route[i])) reached[nr++] = i;
  if (p->compute_closeness) {
    for (uint32_t r = 0; r < nr; r++) {
      uint32_t to = reached[r];
      if (to == src) continue;
      float bsd = t.best_route[to], bs = t.best_secs[to];
      if (!isfinite(bsd) || !isfinite(bs)) continue;
      if (p->track_reach)
        for (uint32_t i = 0; i < T; i++)
          if (bs <= (float)p->seconds[i]) atomic_fetch_add(&jb->reach[i], 1);
      for (uint32_t i = 0; i < T; i++) {
        if (bs <= (float)p->seconds[i]) {
          atomic_add_f64(&OUT(0, i, to), (double)wt);
          float far_ang = p->farness_scaling_offset + (bsd / p->angular_scaling_unit);
          if (g_mutation == OG_MUT_ANG_FAR_SWAP) far_ang = p->angular_scaling_unit + (bsd / p->farness_scaling_offset);
          if (g_mutation == OG_MUT_ANG_FAR_NO_OFFSET) far_ang = bsd / p->angular_scaling_unit;
          atomic_add_f64(&OUT(1, i, to), (double)(far_ang * wt));
          float harm_ang = 1.0f + (bsd / p->angular_scaling_unit);
          atomic_add_f64(&OUT(2, i, to), (double)((1.0f / harm_ang) * wt));
        }
      }
    }
  }
  if (p->compute_betweenness) {
    const uint32_t S = 2 * N;
    uint32_t *sorted = xmalloc(sizeof(uint32_t) * (t.n_order ? t.n_order : 1));
    uint32_t ns = sorted_states(&t, sorted);
    double *seed = xmalloc(sizeof(double) * S), *seed_b = calloc(S, sizeof(double));
    double *delta = xmalloc(sizeof(double) * S), *delta_b = xmalloc(sizeof(double) * S);
    for (uint32_t i = 0; i < T; i++) {
      float sthr = (float)p->seconds[i];
      memset(seed, 0, sizeof(double) * S);
      for (uint32_t r = 0; r < nr; r++) {
        uint32_t to = reached[r];
        if (to == src) continue;
        /* best_angular_target_states: centrality.rs:793-821 */
        float br = t.best_route[to], bs = t.best_secs[to];
        if (!isfinite(br) || !isfinite(bs) || bs > sthr) continue;
        uint32_t best[2];
        uint32_t nb = 0;
        for (uint32_t slot = 0; slot < 2; slot++) {
          uint32_t st = 2 * to + slot;
          if (t.sigma[st] == 0.0 || t.secs[st] > sthr) continue;
          if (t.route[st] <= br * (1.0f + p->tolerance)) best[nb++] = st;
        }
        if (nb == 0) continue;
        double pc = p->eligible[to] ? 0.5 : 1.0;
        if (g_mutation == OG_MUT_ANG_FULL_SEED) pc = 1.0;
        double total = 0.0;
        for (uint32_t j = 0; j < nb; j++) total += t.sigma[best[j]];
        if (total == 0.0) continue;
        for (uint32_t j = 0; j < nb; j++)
          seed[best[j]] += pc * (g_mutation == OG_MUT_ANG_NO_SPLIT ? 1.0 : t.sigma[best[j]] / total);
      }
      backprop(&t, sorted, ns, src, seed, seed_b, 1, sthr, 2, delta, delta_b, &OUT(3, i, 0), NULL,
               (double)wt);
    }
    free(sorted);
    free(seed);
    free(seed_b);
    free(delta);
    free(delta_b);
  }
  free(reached);
  trav_free(&t);
}

/* ------------------------------------------------------------------------- */
/* single-predecessor trees                                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
  uint8_t visited, discovered;
  int32_t pred;
  float short_dist, simpl_dist, agg_seconds;
  int32_t origin_seg, last_seg;
} nvisit; /* graph.rs NodeVisit */

typedef struct {
  uint8_t visited;
  int32_t start, end;
  int64_t edge_key;
} evisit; /* graph.rs EdgeVisit */

static void nvisit_init(nvisit *v, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) {
    v[i].visited = v[i].discovered = 0;
    v[i].pred = -1;
    v[i].short_dist = v[i].simpl_dist = v[i].agg_seconds = INFINITY;
    v[i].origin_seg = v[i].last_seg = -1;
  }
}

/* dijkstra_tree_shortest_inner: centrality.rs:1141-1200 */
static uint32_t tree_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                              nvisit *tm, uint32_t *order) {
  nvisit_init(tm, g->n);
  uint32_t no = 0;
  tm[src].agg_seconds = 0.0f;
  tm[src].discovered = 1;
  tm[src].short_dist = 0.0f;
  heap h = {0};
  hpush(&h, 0.0f, 0.0f, src);
  hentry top;
  const float maxsec = (float)max_seconds;
  while (hpop(&h, &top)) {
    uint32_t u = top.idx;
    if (tm[u].visited) continue;
    tm[u].visited = 1;
    order[no++] = u;
    for (int64_t e = g->in_first[u]; e >= 0; e = g->in_next[e]) {
      uint32_t nb = g->esrc[e];
      if (nb == u || tm[nb].visited) continue;
      if (tm[u].pred >= 0 && (int32_t)nb == tm[u].pred) continue;
      float es = edge_travel_seconds(g, nb, u, (uint32_t)e, speed, 1);
      float ts = tm[u].agg_seconds + es;
      if (ts > maxsec) continue;
      if (ts < tm[nb].agg_seconds) {
        tm[nb].short_dist = ts * speed;
        tm[nb].agg_seconds = ts;
        tm[nb].pred = (int32_t)u;
        tm[nb].discovered = 1;
        hpush(&h, ts, ts, nb);
      }
    }
  }
  free(h.a);
  return no;
}

/* dijkstra_tree_angular: centrality.rs:1202-1332 */
static uint32_t tree_simplest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                              nvisit *tm, uint32_t *order) {
  const uint32_t N = g->n, S = 2 * N;
  nvisit_init(tm, N);
  uint8_t *svis = calloc(S, 1);
  float *smetric = xmalloc(sizeof(float) * S), *ssimpl = xmalloc(sizeof(float) * S),
        *ssecs = xmalloc(sizeof(float) * S);
  for (uint32_t i = 0; i < S; i++) smetric[i] = ssimpl[i] = ssecs[i] = INFINITY;
  uint8_t *reached = calloc(N, 1);
  uint32_t no = 0;
  order[no++] = src;
  reached[src] = 1;
  tm[src].discovered = 1;
  tm[src].visited = 1;
  tm[src].simpl_dist = 0.0f;
  tm[src].agg_seconds = 0.0f;
  heap h = {0};
  for (uint32_t slot = 0; slot < 2; slot++) {
    uint32_t s = 2 * src + slot;
    smetric[s] = 0.0f;
    ssimpl[s] = 0.0f;
    ssecs[s] = 0.0f;
    hpush(&h, 0.0f, 0.0f, s);
  }
  const float maxsec = (float)max_seconds;
  hentry top;
  while (hpop(&h, &top)) {
    uint32_t us = top.idx;
    if (svis[us]) continue;
    svis[us] = 1;
    uint32_t u = us / 2, entry = us % 2;
    for (int64_t e = g->out_first[u]; e >= 0; e = g->out_next[e]) {
      uint32_t v = g->edst[e];
      int32_t key = g->epk[e];
      if (slot_pos(g, u, key) != (int)(1 - entry)) continue;
      uint32_t vs = 2 * v + (uint32_t)slot_pos(g, v, key);
      float es = edge_travel_seconds(g, u, v, (uint32_t)e, speed, 0);
      float cs = ssecs[us] + es;
      if (cs > maxsec) continue;
      float csimpl = ssimpl[us] + g->eang[e];
      float cmetric = csimpl + ANGULAR_ROUTE_TIE_BREAK_FACTOR * g->elen[e];
      int improved = cmetric + TIE_EPSILON < smetric[vs];
      int tied = fabsf(cmetric - smetric[vs]) <= TIE_EPSILON;
      if (improved || (tied && cs < ssecs[vs] && us != vs)) {
        smetric[vs] = cmetric;
        ssimpl[vs] = csimpl;
        ssecs[vs] = cs;
        hpush(&h, cmetric, cs, vs);
        nvisit *nv = &tm[v];
        if (!reached[v]) {
          reached[v] = 1;
          order[no++] = v;
        }
        int nimp = csimpl + TIE_EPSILON < nv->simpl_dist;
        int ntied = fabsf(csimpl - nv->simpl_dist) <= TIE_EPSILON;
        if (!nv->discovered || nimp || (ntied && cs < nv->agg_seconds)) {
          nv->discovered = 1;
          nv->visited = 1;
          nv->simpl_dist = csimpl;
          nv->agg_seconds = cs;
          nv->pred = (int32_t)u;
        }
      }
    }
  }
  free(h.a);
  free(svis);
  free(smetric);
  free(ssimpl);
  free(ssecs);
  free(reached);
  return no;
}

/* dijkstra_tree_segment: centrality.rs:1524-1611 */
static uint32_t tree_segment(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                             nvisit *tm, evisit *em, uint32_t *order, uint32_t *eorder,
                             uint32_t *n_eorder) {
  nvisit_init(tm, g->n);
  for (uint32_t e = 0; e < g->m; e++) {
    em[e].visited = 0;
    em[e].start = em[e].end = -1;
    em[e].edge_key = -1;
  }
  uint32_t no = 0, ne = 0;
  tm[src].short_dist = 0.0f;
  tm[src].agg_seconds = 0.0f;
  tm[src].discovered = 1;
  heap h = {0};
  hpush(&h, 0.0f, 0.0f, src);
  const float maxsec = (float)max_seconds;
  hentry top;
  while (hpop(&h, &top)) {
    uint32_t u = top.idx;
    if (tm[u].visited) continue;
    tm[u].visited = 1;
    order[no++] = u;
    for (int64_t e = g->in_first[u]; e >= 0; e = g->in_next[e]) {
      uint32_t nb = g->esrc[e];
      if (nb == u) {
        eorder[ne++] = (uint32_t)e;
        em[e].visited = 1;
        em[e].start = (int32_t)u;
        em[e].end = (int32_t)nb;
        em[e].edge_key = g->ekey[e];
        continue;
      }
      if (tm[nb].visited) continue;
      eorder[ne++] = (uint32_t)e;
      em[e].visited = 1;
      em[e].start = (int32_t)u;
      em[e].end = (int32_t)nb;
      em[e].edge_key = g->ekey[e];
      float es = edge_travel_seconds(g, nb, u, (uint32_t)e, speed, 1);
      float ts = tm[u].agg_seconds + es;
      if (ts > maxsec) continue;
      if (ts < tm[nb].agg_seconds) {
        int32_t origin = (u == src) ? (int32_t)e : tm[u].origin_seg;
        tm[nb].short_dist = ts * speed;
        tm[nb].agg_seconds = ts;
        tm[nb].pred = (int32_t)u;
        tm[nb].origin_seg = origin;
        tm[nb].last_seg = (int32_t)e;
        tm[nb].discovered = 1;
        hpush(&h, ts, ts, nb);
      }
    }
  }
  free(h.a);
  *n_eorder = ne;
  return no;
}

/* segment_centrality per-source body: centrality.rs:2190-2403 */
static void segment_source(job *jb, uint32_t s_i) {
  const og_graph *g = jb->g;
  const og_params *p = jb->p;
  double *out = jb->out;
  const uint32_t N = g->n, T = p->n_thresholds;
  uint32_t src = p->sources[s_i];
  if (!g->live[src]) return;
  uint32_t maxs = max_u32(p->seconds, T);
  nvisit *tm = xmalloc(sizeof(nvisit) * N);
  evisit *em = xmalloc(sizeof(evisit) * (g->m ? g->m : 1));
  uint32_t *order = xmalloc(sizeof(uint32_t) * N);
  uint32_t *eorder = xmalloc(sizeof(uint32_t) * (g->m ? g->m : 1));
  uint32_t ne = 0;
  uint32_t no = tree_segment(g, src, maxs, p->speed_m_s, tm, em, order, eorder, &ne);
  for (uint32_t k = 0; k < ne && p->compute_closeness; k++) {
    const evisit *ev = &em[eorder[k]];
    uint32_t sn = (uint32_t)ev->start, en = (uint32_t)ev->end;
    const nvisit *vn = &tm[sn], *vm = &tm[en];
    if (!isfinite(vn->short_dist) && !isfinite(vm->short_dist)) continue;
    int n_nearer = vn->short_dist <= vm->short_dist;
    float a = n_nearer ? vn->short_dist : vm->short_dist;
    float a_imp = a;
    float b = n_nearer ? vm->short_dist : vn->short_dist;
    float b_imp = b;
    int64_t le = edge_connecting(g, sn, en, (uint32_t)ev->edge_key);
    if (le < 0) {
      atomic_store(&jb->err, OG_ENOEDGE);
      break;
    }
    float elen = g->elen[le], imp = g->eimp[le];
    float c = (elen + a + b) / 2.0f;
    float d = c;
    float c_imp = a_imp + (c - a) * imp;
    float d_imp = c_imp;
    for (int i = (int)T - 1; i >= 0; i--) {
      float dist = (float)p->distances[i];
      float beta = p->betas[i];
      float nbeta = -beta;
      float inv = beta != 0.0f ? 1.0f / nbeta : 0.0f;
      if (a < dist) {
        float cc = c, cci = c_imp;
        if (cc > dist) {
          cc = dist;
          cci = a_imp + (dist - a) * imp;
        }
        atomic_add_f64(&OUT(0, i, src), (double)(cc - a));
        float sh = a_imp < 1.0f ? logf(cci) : logf(fmaxf(cci / a_imp, 1.1920929e-07f));
        atomic_add_f64(&OUT(1, i, src), (double)sh);
        float bet = beta == 0.0f ? cci - a_imp : (expf(nbeta * cci) - expf(nbeta * a_imp)) * inv;
        atomic_add_f64(&OUT(2, i, src), (double)bet);
      }
      if (b == d) continue;
      if (b <= dist) {
        float dd = d, ddi = d_imp;
        if (dd > dist) {
          dd = dist;
          ddi = b_imp + (dist - b) * imp;
        }
        atomic_add_f64(&OUT(0, i, src), (double)(dd - b));
        float sh = b_imp < 1.0f ? logf(ddi) : logf(fmaxf(ddi / b_imp, 1.1920929e-07f));
        atomic_add_f64(&OUT(1, i, src), (double)sh);
        float bet = beta == 0.0f ? ddi - b_imp : (expf(nbeta * ddi) - expf(nbeta * b_imp)) * inv;
        atomic_add_f64(&OUT(2, i, src), (double)bet);
      }
    }
  }
  for (uint32_t k = 0; k < no && p->compute_betweenness; k++) {
    uint32_t to = order[k];
    if (to <= src) continue;
    const nvisit *tv = &tm[to];
    if (!isfinite(tv->short_dist)) continue;
    const evisit *oe = &em[tv->origin_seg], *le = &em[tv->last_seg];
    int64_t oi = edge_connecting(g, (uint32_t)oe->start, (uint32_t)oe->end, (uint32_t)oe->edge_key);
    int64_t li = edge_connecting(g, (uint32_t)le->start, (uint32_t)le->end, (uint32_t)le->edge_key);
    if (oi < 0 || li < 0) {
      atomic_store(&jb->err, OG_ENOEDGE);
      break;
    }
    float o_len = g->elen[oi], l_len = g->elen[li];
    float min_span = tv->short_dist - o_len - l_len;
    float o_1 = min_span, o_2 = min_span + o_len, l_1 = min_span, l_2 = min_span + l_len;
    int32_t inter = tv->pred;
    while (inter >= 0) {
      if ((uint32_t)inter == src) break;
      for (int i = (int)T - 1; i >= 0; i--) {
        float dist = (float)p->distances[i];
        float beta = p->betas[i];
        if (min_span <= dist) {
          float o2s = fminf(o_2, dist), l2s = fminf(l_2, dist);
          float auc;
          if (beta == 0.0f)
            auc = (o2s - o_1) + (l2s - l_1);
          else {
            float nb = -beta, inv = 1.0f / nb;
            auc = (expf(nb * o2s) - expf(nb * o_1)) * inv + (expf(nb * l2s) - expf(nb * l_1)) * inv;
          }
          if (isfinite(auc) && auc >= 0.0f) atomic_add_f64(&OUT(3, i, inter), (double)auc);
        }
      }
      inter = tm[inter].pred;
    }
  }
  free(tm);
  free(em);
  free(order);
  free(eorder);
}

static void *worker(void *arg) {
  job *jb = (job *)arg;
  for (;;) {
    uint32_t i = atomic_fetch_add(&jb->next, 1);
    if (i >= jb->p->n_sources) break;
    if (jb->kind == 0)
      shortest_source(jb, i);
    else if (jb->kind == 1)
      simplest_source(jb, i);
    else if (jb->kind == 2)
      segment_source(jb, i);
    else
      od_source(jb, i);
  }
  return NULL;
}

static int run_job_od(const og_graph *g, const og_params *p, double *out, uint32_t *reach, int kind,
                      uint32_t M, const uint32_t *od_off, const uint32_t *od_dest, const float *od_w);

static int run_job(const og_graph *g, const og_params *p, double *out, uint32_t *reach, int kind,
                   uint32_t M) {
  return run_job_od(g, p, out, reach, kind, M, NULL, NULL, NULL);
}

static int run_job_od(const og_graph *g, const og_params *p, double *out, uint32_t *reach, int kind,
                      uint32_t M, const uint32_t *od_off, const uint32_t *od_dest, const float *od_w) {
  const uint32_t N = g->n, T = p->n_thresholds;
  if (T == 0 || T > OG_MAX_T) return OG_EINVAL;
  memset(out, 0, sizeof(double) * (size_t)M * T * N);
  _Atomic uint32_t areach[OG_MAX_T];
  for (uint32_t i = 0; i < OG_MAX_T; i++) atomic_init(&areach[i], 0);
  job jb;
  jb.g = g;
  jb.p = p;
  jb.out = out;
  jb.reach = areach;
  atomic_init(&jb.next, 0);
  atomic_init(&jb.err, OG_OK);
  jb.kind = kind;
  jb.od_off = od_off;
  jb.od_dest = od_dest;
  jb.od_w = od_w;
  int nt = p->n_threads > 0 ? p->n_threads : 1;
  if (nt == 1) {
    worker(&jb);
  } else {
    pthread_t *th = xmalloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, worker, &jb);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
  }
  if (reach)
    for (uint32_t i = 0; i < T; i++) reach[i] = atomic_load(&areach[i]);
  return atomic_load(&jb.err);
}

/* scale_metric_results: centrality.rs:875-893 */
static void scale_rows(double *out, uint32_t N, uint32_t T, uint32_t m0, uint32_t m1, double s) {
  if (s == 1.0) return;
  for (uint32_t m = m0; m <= m1; m++)
    for (uint32_t i = 0; i < T; i++)
      for (uint32_t v = 0; v < N; v++) out[((size_t)m * T + i) * N + v] *= s;
}

int og_centrality_shortest(const og_graph *g, const og_params *p, double *out, uint32_t *reach) {
  int rc = run_job(g, p, out, reach, 0, 7);
  if (rc == OG_OK && p->compute_betweenness) scale_rows(out, g->n, p->n_thresholds, 5, 6, p->betw_scale);
  return rc;
}

int og_betweenness_od_shortest(const og_graph *g, const og_params *p, const uint32_t *od_off,
                               const uint32_t *od_dest, const float *od_w, double *out) {
  for (uint32_t i = 0; i < p->n_sources; i++)
    for (uint32_t k = od_off[i]; k < od_off[i + 1]; k++)
      if (od_dest[k] >= g->n || p->sources[i] >= g->n) return OG_EINVAL;
  return run_job_od(g, p, out, NULL, 3, 2, od_off, od_dest, od_w);
}

int og_centrality_simplest(const og_graph *g, const og_params *p, double *out, uint32_t *reach) {
  if (!g->is_dual || !g->slots_ok) return OG_EINVAL;
  int rc = run_job(g, p, out, reach, 1, 4);
  if (rc == OG_OK && p->compute_betweenness) scale_rows(out, g->n, p->n_thresholds, 3, 3, p->betw_scale);
  return rc;
}

int og_segment_centrality(const og_graph *g, const og_params *p, double *out) {
  return run_job(g, p, out, NULL, 2, 4);
}

/* ------------------------------------------------------------------------- */
/* single-source exports                                                     */
/* ------------------------------------------------------------------------- */

static void export_trav(const traversal *t, float *secs, float *route, double *sigma,
                        uint32_t *n_preds, uint32_t *preds, uint32_t max_preds, uint32_t *order,
                        uint32_t *n_order) {
  for (uint32_t s = 0; s < t->n_states; s++) {
    secs[s] = t->secs[s];
    route[s] = t->route[s];
    sigma[s] = t->sigma[s];
    n_preds[s] = t->preds[s].n;
    for (uint32_t j = 0; j < t->preds[s].n && j < max_preds; j++)
      preds[(size_t)s * max_preds + j] = t->preds[s].v[j];
  }
  for (uint32_t i = 0; i < t->n_order; i++) order[i] = t->order[i];
  *n_order = t->n_order;
}

int og_brandes_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                        float tolerance, float *secs, float *route, double *sigma,
                        uint32_t *n_preds, uint32_t *preds, uint32_t max_preds, uint32_t *order,
                        uint32_t *n_order) {
  if (src >= g->n) return OG_EINVAL;
  traversal t;
  brandes_shortest(g, src, max_seconds, speed, tolerance, &t);
  export_trav(&t, secs, route, sigma, n_preds, preds, max_preds, order, n_order);
  trav_free(&t);
  return OG_OK;
}

int og_brandes_angular(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                       float tolerance, float *secs, float *route, double *sigma,
                       uint32_t *n_preds, uint32_t *preds, uint32_t max_preds, uint32_t *order,
                       uint32_t *n_order, float *best_route, float *best_secs) {
  if (src >= g->n || !g->is_dual || !g->slots_ok) return OG_EINVAL;
  traversal t;
  brandes_angular(g, src, max_seconds, speed, tolerance, &t);
  export_trav(&t, secs, route, sigma, n_preds, preds, max_preds, order, n_order);
  for (uint32_t i = 0; i < g->n; i++) {
    best_route[i] = t.best_route[i];
    best_secs[i] = t.best_secs[i];
  }
  trav_free(&t);
  return OG_OK;
}

static void export_tree(const nvisit *tm, uint32_t n, uint8_t *visited, uint8_t *discovered,
                        int32_t *pred, float *dist, int use_simpl, float *agg_seconds) {
  for (uint32_t i = 0; i < n; i++) {
    visited[i] = tm[i].visited;
    discovered[i] = tm[i].discovered;
    pred[i] = tm[i].pred;
    dist[i] = use_simpl ? tm[i].simpl_dist : tm[i].short_dist;
    agg_seconds[i] = tm[i].agg_seconds;
  }
}

int og_tree_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                     uint8_t *visited, uint8_t *discovered, int32_t *pred, float *short_dist,
                     float *agg_seconds, uint32_t *order, uint32_t *n_order) {
  if (src >= g->n || !(speed > 0.0f) || !isfinite(speed)) return OG_EINVAL;
  nvisit *tm = xmalloc(sizeof(nvisit) * g->n);
  *n_order = tree_shortest(g, src, max_seconds, speed, tm, order);
  export_tree(tm, g->n, visited, discovered, pred, short_dist, 0, agg_seconds);
  free(tm);
  return OG_OK;
}

int og_tree_simplest(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                     uint8_t *visited, uint8_t *discovered, int32_t *pred, float *simpl_dist,
                     float *agg_seconds, uint32_t *order, uint32_t *n_order) {
  if (!g->is_dual || !g->slots_ok) return OG_EINVAL;
  if (src >= g->n || !(speed > 0.0f) || !isfinite(speed)) return OG_EINVAL;
  nvisit *tm = xmalloc(sizeof(nvisit) * g->n);
  *n_order = tree_simplest(g, src, max_seconds, speed, tm, order);
  export_tree(tm, g->n, visited, discovered, pred, simpl_dist, 1, agg_seconds);
  free(tm);
  return OG_OK;
}

int og_tree_segment(const og_graph *g, uint32_t src, uint32_t max_seconds, float speed,
                    uint8_t *visited, uint8_t *discovered, int32_t *pred, float *short_dist,
                    float *agg_seconds, int32_t *origin_seg, int32_t *last_seg, uint32_t *order,
                    uint32_t *n_order, uint32_t *edge_order, uint32_t *n_edges) {
  if (src >= g->n || !(speed > 0.0f) || !isfinite(speed)) return OG_EINVAL;
  nvisit *tm = xmalloc(sizeof(nvisit) * g->n);
  evisit *em = xmalloc(sizeof(evisit) * (g->m ? g->m : 1));
  *n_order = tree_segment(g, src, max_seconds, speed, tm, em, order, edge_order, n_edges);
  export_tree(tm, g->n, visited, discovered, pred, short_dist, 0, agg_seconds);
  for (uint32_t i = 0; i < g->n; i++) {
    origin_seg[i] = tm[i].origin_seg;
    last_seg[i] = tm[i].last_seg;
  }
  free(tm);
  free(em);
  return OG_OK;
}

/* ------------------------------------------------------------------------- */
/* land-use accessibility, threaded (DataMap.accessibility)                  */
/* ------------------------------------------------------------------------- */

/* Per live source: its shortest-path tree (simplest-path with `angular`),
 * then aggregate_to_src_idx (data.rs:520-615): nodes in visit order below the
 * walk limit, each node's members in assignment order, one entry per dedupe
 * group kept, replaced only by a strictly smaller total; then the counts of
 * DataMap.accessibility (data.rs:680-741). Members: n_assign triples (node,
 * entry, distance f64, used as f32); entry_group[e] its dedupe group (<
 * n_entries); entry_class[e] its class (-1: not counted). Each source writes
 * its own rows, so threads need no atomics. out_count, out_weighted:
 * [n_classes][T][N]; out_nearest: [n_classes][N] (NaN: none). */
typedef struct {
  const og_graph *g;
  const uint32_t *node_off, *node_mem, *assign_entry;
  const float *assign_dist;
  const uint32_t *entry_group;
  const int32_t *entry_class;
  uint32_t n_entries, n_classes, T, max_seconds;
  const uint32_t *distances;
  const float *betas, *mcws;
  float speed;
  int angular;
  const uint8_t *live;
  double *count, *weighted, *nearest;
  _Atomic uint32_t next;
} acc_job;

static void *acc_worker(void *arg) {
  acc_job *jb = (acc_job *)arg;
  const og_graph *g = jb->g;
  const uint32_t N = g->n, T = jb->T;
  nvisit *tm = xmalloc(sizeof(nvisit) * N);
  uint32_t *order = xmalloc(sizeof(uint32_t) * (N ? N : 1));
  uint32_t *tag = calloc(jb->n_entries ? jb->n_entries : 1, sizeof(uint32_t));
  uint32_t *win = xmalloc(sizeof(uint32_t) * (jb->n_entries ? jb->n_entries : 1));
  float *wtot = xmalloc(sizeof(float) * (jb->n_entries ? jb->n_entries : 1));
  uint32_t *touched = xmalloc(sizeof(uint32_t) * (jb->n_entries ? jb->n_entries : 1));
  const float maxsec = (float)jb->max_seconds;
  const float max_walk = maxsec * jb->speed;
  uint32_t dmax = 0;
  for (uint32_t t = 0; t < T; t++) dmax = jb->distances[t] > dmax ? jb->distances[t] : dmax;
  uint32_t epoch = 0;
  for (;;) {
    const uint32_t src = atomic_fetch_add(&jb->next, 1);
    if (src >= N) break;
    if (!jb->live[src]) continue;
    const uint32_t no = jb->angular ? tree_simplest(g, src, jb->max_seconds, jb->speed, tm, order)
                                    : tree_shortest(g, src, jb->max_seconds, jb->speed, tm, order);
    ++epoch;
    uint32_t nt = 0;
    for (uint32_t k = 0; k < no; k++) {
      const uint32_t v = order[k];
      const float agg = tm[v].agg_seconds;
      if (agg >= maxsec) continue;
      const float net = agg * jb->speed;
      for (uint32_t q = jb->node_off[v]; q < jb->node_off[v + 1]; q++) {
        const uint32_t m = jb->node_mem[q];
        const uint32_t e = jb->assign_entry[m];
        const float total = net + jb->assign_dist[m];
        if (!(total <= max_walk)) continue;
        const uint32_t grp = jb->entry_group[e];
        if (tag[grp] != epoch) {
          tag[grp] = epoch;
          win[grp] = e;
          wtot[grp] = total;
          touched[nt++] = grp;
        } else if (total < wtot[grp]) {
          win[grp] = e;
          wtot[grp] = total;
        }
      }
    }
    for (uint32_t k = 0; k < nt; k++) {
      const uint32_t grp = touched[k];
      const int32_t c = jb->entry_class[win[grp]];
      if (c < 0) continue;
      const float d = wtot[grp];
      for (uint32_t t = 0; t < T; t++) {
        if (!(d <= (float)jb->distances[t])) continue;
        const size_t row = ((size_t)c * T + t) * N + src;
        jb->count[row] += 1.0;
        /* clipped_beta_wt (common.rs:339-348) */
        const float raw = expf(-jb->betas[t] * d);
        jb->weighted[row] += (double)(fminf(raw, jb->mcws[t]) / jb->mcws[t]);
        if (jb->distances[t] == dmax) {
          double *nr = &jb->nearest[(size_t)c * N + src];
          if (isnan(*nr) || (double)d < *nr) *nr = (double)d;
        }
      }
    }
  }
  free(tm);
  free(order);
  free(tag);
  free(win);
  free(wtot);
  free(touched);
  return NULL;
}

int og_accessibility(const og_graph *g, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                     uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node,
                     const uint32_t *assign_entry, const double *assign_dist, uint32_t T, const uint32_t *distances,
                     const float *betas, const float *mcws, uint32_t max_seconds, float speed, int angular,
                     const uint8_t *live, int n_threads, double *out_count, double *out_weighted,
                     double *out_nearest) {
  const uint32_t N = g->n;
  if (T == 0 || !(speed > 0.0f) || !isfinite(speed)) return OG_EINVAL;
  if (angular && (!g->is_dual || !g->slots_ok)) return OG_EINVAL;
  for (uint32_t e = 0; e < n_entries; e++) {
    if (entry_group[e] >= n_entries) return OG_EINVAL;
    if (entry_class[e] >= (int32_t)n_classes) return OG_EINVAL;
  }
  for (uint32_t i = 0; i < n_assign; i++)
    if (assign_node[i] >= N || assign_entry[i] >= n_entries) return OG_EINVAL;
  for (size_t i = 0; i < (size_t)n_classes * T * N; i++) out_count[i] = out_weighted[i] = 0.0;
  for (size_t i = 0; i < (size_t)n_classes * N; i++) out_nearest[i] = NAN;
  /* members of each node, in assignment order */
  uint32_t *node_off = calloc((size_t)N + 1, sizeof(uint32_t));
  uint32_t *node_mem = xmalloc(sizeof(uint32_t) * (n_assign ? n_assign : 1));
  float *adist = xmalloc(sizeof(float) * (n_assign ? n_assign : 1));
  for (uint32_t i = 0; i < n_assign; i++) {
    node_off[assign_node[i] + 1]++;
    adist[i] = (float)assign_dist[i]; /* data.rs:548 casts the f64 distance to f32 */
  }
  for (uint32_t v = 0; v < N; v++) node_off[v + 1] += node_off[v];
  {
    uint32_t *at = xmalloc(sizeof(uint32_t) * (N ? N : 1));
    for (uint32_t v = 0; v < N; v++) at[v] = node_off[v];
    for (uint32_t i = 0; i < n_assign; i++) node_mem[at[assign_node[i]]++] = i;
    free(at);
  }
  acc_job jb;
  jb.g = g;
  jb.node_off = node_off;
  jb.node_mem = node_mem;
  jb.assign_entry = assign_entry;
  jb.assign_dist = adist;
  jb.entry_group = entry_group;
  jb.entry_class = entry_class;
  jb.n_entries = n_entries;
  jb.n_classes = n_classes;
  jb.T = T;
  jb.max_seconds = max_seconds;
  jb.distances = distances;
  jb.betas = betas;
  jb.mcws = mcws;
  jb.speed = speed;
  jb.angular = angular;
  jb.live = live;
  jb.count = out_count;
  jb.weighted = out_weighted;
  jb.nearest = out_nearest;
  atomic_init(&jb.next, 0);
  const int nt = n_threads > 0 ? n_threads : 1;
  if (nt == 1) {
    acc_worker(&jb);
  } else {
    pthread_t *th = xmalloc(sizeof(pthread_t) * nt);
    for (int i = 0; i < nt; i++) pthread_create(&th[i], NULL, acc_worker, &jb);
    for (int i = 0; i < nt; i++) pthread_join(th[i], NULL);
    free(th);
  }
  free(node_off);
  free(node_mem);
  free(adist);
  return OG_OK;
}
