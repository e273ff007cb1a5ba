/*
 * cityseer_oracle.h -- CPU restatement of cityseer's localised network-centrality
 * path (TEST INFRASTRUCTURE ONLY).
 *
 * This library is the parity checker for the HIP implementation and the CPU
 * baseline timed by bench.py. It is never linked into, loaded by or called from
 * the product library. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * Every function restates the reference algorithm it names (file:line into the
 * cityseer repository). Arithmetic is f32 unless marked f64, one IEEE op per
 * source-level operation (compiled with -ffp-contract=off, no fast-math).
 *
 * Equal route costs: the reference's std BinaryHeap pops equal keys in an
 * order its internals decide (rust/src/centrality.rs:358-386). The oracle has
 * two modes (og_set_heap_mode): "rust" runs the std BinaryHeap algorithm and
 * reproduces that order; "keyed" (default) pops equal route costs by
 * (seconds, index), the order the GPU search reproduces. On inputs without
 * exact ties the two modes give bit-identical results.
 */
#ifndef CITYSEER_ORACLE_H
#define CITYSEER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_graph og_graph;

/* Status codes shared by all functions. */
#define OG_OK 0
#define OG_EINVAL 1     /* invalid argument (reference raises ValueError) */
#define OG_ENOEDGE 2    /* edge lookup failed (reference panics "Edge not found") */
#define OG_ENOMEM 3

/* 0 (default): keyed pop order (route, seconds, index); 1: the pop order of
 * Rust's std BinaryHeap (centrality.rs:358-386), for pinning against the
 * reference's golden vectors on graphs with exact ties. Process-wide. */
void og_set_heap_mode(int rust_order);

/* Planted mistakes for the pinning tests (0 = none, the default):
 *  OG_MUT_BETA_SOURCE_ROUTE   the beta seed takes the source's route instead of the target's
 *  OG_MUT_BETA_FULL_SEED      the beta seed drops the 0.5 pair weight
 *  OG_MUT_ANG_FAR_SWAP        angular farness with scaling unit and offset swapped
 *  OG_MUT_ANG_FAR_NO_OFFSET   angular farness without the offset
 *  OG_MUT_ANG_FULL_SEED       angular betweenness seeds drop the 0.5 pair weight
 *  OG_MUT_ANG_NO_SPLIT        angular seeds not split over tied terminal states */
enum {
  OG_MUT_NONE = 0,
  OG_MUT_BETA_SOURCE_ROUTE = 1,
  OG_MUT_BETA_FULL_SEED = 2,
  OG_MUT_ANG_FAR_SWAP = 4,
  OG_MUT_ANG_FAR_NO_OFFSET = 5,
  OG_MUT_ANG_FULL_SEED = 6,
  OG_MUT_ANG_NO_SPLIT = 7,
};
void og_set_mutation(int m);

/* Graph built from nodes and directed edges given in insertion order.
 * z: NaN = node has no elevation. seconds: NaN = derive from length
 * (rust/src/graph.rs:879 stores NaN for street edges). primal_key: -1 = none. */
og_graph *og_graph_new(uint32_t n_nodes, const double *z, const uint8_t *live,
                       const float *weight, uint32_t n_edges, const uint32_t *src,
                       const uint32_t *dst, const uint32_t *edge_key,
                       const float *length, const float *imp_factor,
                       const float *angle_sum, const float *seconds,
                       const int32_t *primal_key, int is_dual);
void og_graph_free(og_graph *g);

/* rust/src/common.rs:241-270. mode 0: distances given, 1: betas, 2: minutes.
 * Writes n thresholds into out_*; returns OG_EINVAL on the reference's errors. */
int og_pair_thresholds(float speed_m_s, int mode, const uint32_t *distances,
                       const float *betas, const float *minutes, uint32_t n,
                       float min_threshold_wt, uint32_t *out_distances,
                       float *out_betas, uint32_t *out_seconds);

typedef struct {
  uint32_t n_thresholds;
  const uint32_t *distances; /* [T] */
  const float *betas;        /* [T] */
  const uint32_t *seconds;   /* [T] */
  float speed_m_s;
  float tolerance;           /* decimal, already clamped >= 1e-4 */
  int compute_closeness;
  int compute_betweenness;
  uint32_t n_sources;
  const uint32_t *sources;   /* [n_sources] node indices */
  const float *source_wt;    /* [n_sources]; NaN = source skipped (sampling) */
  const uint8_t *eligible;   /* [N] source-eligible mask */
  int track_reach;           /* accumulate reachability_totals */
  double betw_scale;         /* post-hoc betweenness scale (centrality.rs:1852-1868) */
  float angular_scaling_unit;
  float farness_scaling_offset;
  int n_threads;             /* 1 = deterministic summation order */
} og_params;

/* Outputs are f64 [M][T][N] with M = 7 for shortest
 * (density, farness, cycles, harmonic, beta, betweenness, betweenness_beta),
 * M = 4 for simplest (density, farness, harmonic, betweenness) and
 * M = 4 for segment (density, harmonic, beta, betweenness). */
int og_centrality_shortest(const og_graph *g, const og_params *p, double *out,
                           uint32_t *reach_totals);
/* betweenness_od_shortest (rust/src/centrality.rs:2413-2540): p->sources are
 * the unique origins, p->n_sources of them; origin i's OD row (destination,
 * trip weight) is od_dest/od_w[od_off[i] .. od_off[i + 1]), destinations
 * unique per origin. out: f64 [2][T][N] = betweenness, betweenness_beta. Only
 * thresholds, speed, tolerance and n_threads of p are read besides sources. */
int og_betweenness_od_shortest(const og_graph *g, const og_params *p, const uint32_t *od_off,
                               const uint32_t *od_dest, const float *od_w, double *out);
int og_centrality_simplest(const og_graph *g, const og_params *p, double *out,
                           uint32_t *reach_totals);
int og_segment_centrality(const og_graph *g, const og_params *p, double *out);

/* Single-source Brandes traversal export (centrality.rs:1344-1494). State
 * arrays are N long. preds are flattened: preds[i*max_preds + j]. */
int og_brandes_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds,
                        float speed_m_s, float tolerance, float *secs,
                        float *route, double *sigma, uint32_t *n_preds,
                        uint32_t *preds, uint32_t max_preds, uint32_t *order,
                        uint32_t *n_order);
/* Angular traversal export (centrality.rs:577-775). State arrays are 2N long,
 * best_* arrays N long. */
int og_brandes_angular(const og_graph *g, uint32_t src, uint32_t max_seconds,
                       float speed_m_s, float tolerance, float *secs,
                       float *route, double *sigma, uint32_t *n_preds,
                       uint32_t *preds, uint32_t max_preds, uint32_t *order,
                       uint32_t *n_order, float *best_route, float *best_secs);

/* Single-predecessor trees (centrality.rs:1141-1332,1524-1611). pred = -1 for
 * none. NodeVisit fields: visited, discovered, pred, short_dist, simpl_dist,
 * agg_seconds, origin_seg, last_seg. */
int og_tree_shortest(const og_graph *g, uint32_t src, uint32_t max_seconds,
                     float speed_m_s, uint8_t *visited, uint8_t *discovered,
                     int32_t *pred, float *short_dist, float *agg_seconds,
                     uint32_t *order, uint32_t *n_order);
int og_tree_simplest(const og_graph *g, uint32_t src, uint32_t max_seconds,
                     float speed_m_s, uint8_t *visited, uint8_t *discovered,
                     int32_t *pred, float *simpl_dist, float *agg_seconds,
                     uint32_t *order, uint32_t *n_order);
int og_tree_segment(const og_graph *g, uint32_t src, uint32_t max_seconds,
                    float speed_m_s, uint8_t *visited, uint8_t *discovered,
                    int32_t *pred, float *short_dist, float *agg_seconds,
                    int32_t *origin_seg, int32_t *last_seg, uint32_t *order,
                    uint32_t *n_order, uint32_t *edge_order, uint32_t *n_edges);


/* Land-use accessibility over all live sources, threaded (data.rs:520-615,
 * 680-741): per source its shortest-path tree (simplest-path with `angular`,
 * dual graphs only), one entry per dedupe group within the walk limit, then
 * per class and threshold the count, the clipped-beta weighted count and, at
 * the largest threshold, the nearest distance. Members are n_assign triples
 * (assign_node, assign_entry, assign_dist). out_count, out_weighted:
 * [n_classes][T][N]; out_nearest: [n_classes][N], NaN where none. */
int og_accessibility(const og_graph *g, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                     uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node,
                     const uint32_t *assign_entry, const double *assign_dist, uint32_t T, const uint32_t *distances,
                     const float *betas, const float *mcws, uint32_t max_seconds, float speed, int angular,
                     const uint8_t *live, int n_threads, double *out_count, double *out_weighted,
                     double *out_nearest);

#ifdef __cplusplus
}
#endif
#endif
