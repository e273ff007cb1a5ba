/*
 * cityseer_amd.h -- C ABI of the MI355X implementation of cityseer's localised
 * network-centrality path.
 *
 * Plain pointers and sizes only. Every compute entry point runs hand-written
 * HIP kernels on the current HIP device; there is no CPU fallback. Functions
 * return CS_OK (0) or a CS_E* status; cs_last_error() returns the calling
 * thread's last message, whose fragments match the reference's ValueError
 * messages where one exists.
 *
 * Each declaration names the reference interface it replaces (file:line in the
 * cityseer repository).
 */
#ifndef CITYSEER_AMD_H
#define CITYSEER_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CS_OK 0
#define CS_EINVAL 1    /* argument error; the Python layer raises ValueError */
#define CS_EDEVICE 2   /* no usable HIP device, or a HIP runtime error */
#define CS_ENOMEM 3    /* device memory exhausted (or HBM budget exceeded) */
#define CS_EINTERNAL 4 /* internal invariant violated */
#define CS_EBUSY 5     /* serving queue full / request not finished */
#define CS_ETIMEDOUT 6 /* cs_server_wait_for: the request did not finish in time */
#define CS_ECANCELED 7 /* the request was cancelled (cs_server_cancel) */

/* Last error message of the calling thread (empty string if none). */
const char *cs_last_error(void);

/* Library version string. */
const char *cs_version(void);

/* ---------------------------------------------------------------------- */
/* Threshold conversion                                                   */
/* ---------------------------------------------------------------------- */

/* Replaces rustalgos.pair_distances_betas_time (rust/src/common.rs:241-270).
 * mode 0: distances given; 1: betas given; 2: minutes given. Each list has n
 * entries; outputs receive n distances, betas and seconds. */
int cs_pair_distances_betas_time(float speed_m_s, int mode, const uint32_t *distances,
                                 const float *betas, const float *minutes, uint32_t n,
                                 float min_threshold_wt, uint32_t *out_distances,
                                 float *out_betas, uint32_t *out_seconds);

/* Replaces rustalgos.distances_from_betas (rust/src/common.rs:98-131). */
int cs_distances_from_betas(const float *betas, uint32_t n, float min_threshold_wt,
                            uint32_t *out_distances);
/* Replaces rustalgos.betas_from_distances (rust/src/common.rs:133-165). */
int cs_betas_from_distances(const uint32_t *distances, uint32_t n, float min_threshold_wt,
                            float *out_betas);
/* Replaces rustalgos.distances_from_seconds (rust/src/common.rs:167-201). */
int cs_distances_from_seconds(const uint32_t *seconds, uint32_t n, float speed_m_s,
                              uint32_t *out_distances);
/* Replaces rustalgos.seconds_from_distances (rust/src/common.rs:203-237). */
int cs_seconds_from_distances(const uint32_t *distances, uint32_t n, float speed_m_s,
                              uint32_t *out_seconds);

/* ---------------------------------------------------------------------- */
/* Graph                                                                  */
/* ---------------------------------------------------------------------- */

typedef struct cs_graph cs_graph;

/* Replaces the NetworkStructure storage built by add_street_node /
 * add_street_edge (rust/src/graph.rs:384-452,728-889). Nodes 0..n_nodes-1;
 * directed edges in insertion order (edge id = position). z: NaN = no
 * elevation (NULL = none). seconds: NaN = derive from length (NULL = all).
 * primal_key: dual edges' shared primal node id, -1 = none (NULL = none).
 * imp_factor / angle_sum NULL = 1.0 / 0.0. The handle is immutable after
 * creation; concurrent calls on one handle are safe. */
int cs_graph_create(uint32_t n_nodes, const double *z, const uint8_t *live, const float *weight,
                    uint32_t n_edges, const uint32_t *src, const uint32_t *dst,
                    const uint32_t *edge_key, const float *length, const float *imp_factor,
                    const float *angle_sum, const float *seconds, const int32_t *primal_key,
                    int is_dual, cs_graph **out);
void cs_graph_destroy(cs_graph *g);

/* Node coordinates (x, y per node, the reference's NetworkStructure node
 * positions, rust/src/graph.rs:426-452). They set the order nodes take in device
 * memory: along a Hilbert curve, so that nodes near in space are near in
 * memory and a shortest-path search touches fewer cache lines. Results are
 * unchanged: every entry point takes and returns original node indices, and
 * exact ties still resolve by original index. Dual graphs, graphs with
 * non-finite coordinates and graphs of fewer than two nodes keep their order.
 * Must come before the graph's first computation (CS_EINVAL after it). No
 * reference counterpart: the reference keeps nodes in insertion order. */
int cs_graph_set_xy(cs_graph *g, const double *x, const double *y);

/* order[d] = original index of the node at device position d (n_nodes
 * entries). Sources listed in this order (cs_centrality_shortest_device) run
 * fastest: neighbouring searches then run side by side. No reference
 * counterpart. */
int cs_graph_node_order(const cs_graph *g, uint32_t *order);

/* Uploads the graph's device arrays (adjacency, node order) to HIP device
 * `device` and, when speed_m_s > 0, its edge travel seconds for that speed,
 * so that the first computation there pays no upload (computations otherwise
 * upload on first use; a graph keeps one copy per device until destroyed).
 * Freezes the device node order (cs_graph_set_xy fails afterwards). No
 * reference counterpart: the reference has no device. */
int cs_graph_upload(cs_graph *g, int device, float speed_m_s);

/* ---------------------------------------------------------------------- */
/* Ingest (host only, no GPU)                                             */
/* ---------------------------------------------------------------------- */

/* The payload NetworkStructure.add_street_edge derives from each edge's
 * polyline (rust/src/graph.rs:764-857), for n_edges edges at once: edge e's
 * points are xy[coord_off[e] .. coord_off[e+1]) ([n][2] f64, at least 2).
 * length = f32 of the summed segment lengths; angle_sum = f32 of the summed
 * turn angles at interior vertices (degrees); in/out bearing = degrees along
 * the first / last distinct coordinate pair (0 when there is none). Any
 * output may be NULL. */
int cs_edge_geometry(uint32_t n_edges, const uint32_t *coord_off, const double *xy, float *length,
                     float *angle_sum, float *in_bearing, float *out_bearing);

/* Splits each polyline into ceil(length / decompose_max) pieces of equal
 * length, as tools.graphs.nx_decompose cuts an edge
 * (pysrc/cityseer/tools/graphs.py:1911-1961, shapely substring and
 * interpolation semantics). pts: [n][dims] f64 with dims 2 (x, y) or 3
 * (x, y, z: z interpolated alike). *n_pieces, *n_points size the result;
 * cs_pieces_copy fills edge_cuts [n_edges], piece_off [n_pieces + 1] and pts
 * [n_points][dims] (any may be NULL); cs_pieces_free releases it. Each
 * interior piece end is a new node. */
typedef struct cs_pieces cs_pieces;
int cs_decompose(uint32_t n_edges, const uint32_t *coord_off, const double *pts, uint32_t dims,
                 double decompose_max, cs_pieces **out, uint32_t *n_pieces, uint32_t *n_points);
/* The two halves of each polyline (cut at half its length, same rules), as
 * tools.graphs.nx_to_dual splits primal edges (graphs.py:1965-2149). */
int cs_halve(uint32_t n_edges, const uint32_t *coord_off, const double *pts, uint32_t dims, cs_pieces **out,
             uint32_t *n_pieces, uint32_t *n_points);
int cs_pieces_copy(const cs_pieces *p, uint32_t *edge_cuts, uint32_t *piece_off, double *pts);
void cs_pieces_free(cs_pieces *p);

/* Geometries in flat arrays: geometry g has parts [geom_part_off[g],
 * geom_part_off[g + 1]); part p has kind part_kind[p] (0 point, 1 line string,
 * 2 polygon) and rings [part_ring_off[p], part_ring_off[p + 1]) (one ring for
 * points and lines; a polygon's exterior, then its holes); ring r has the
 * coordinates [ring_off[r], ring_off[r + 1]) of xy [n][2]. Multi-geometries
 * are several parts. */
typedef struct {
  uint32_t n_geoms;
  const uint32_t *geom_part_off;
  const uint8_t *part_kind;
  const uint32_t *part_ring_off;
  const uint32_t *ring_off;
  const double *xy;
} cs_geoms;

/* Replaces DataMap.assign_data_to_network (rust/src/data.rs:455-512) with
 * NetworkStructure.find_assignments_for_entry (rust/src/graph.rs:1346-1535):
 * each data geometry is assigned to the end nodes of its nearest street edges
 * whose line of sight (from the geometry's nearest point to the node) crosses
 * no barrier and no street. Street edges are given in edge order with their
 * polylines (edge_off [n_edges + 1] into edge_xy [n][2]); edges repeating an
 * earlier edge's node pair and geometry (either direction) are skipped, as
 * the reference's edge R-tree skips them. barriers may be NULL. The result
 * rows (data index, node, distance) come out in data order through
 * cs_assignment_copy; cs_assignment_free releases them. Host only. */
typedef struct cs_assignment cs_assignment;
int cs_assign_data(uint32_t n_nodes, const double *node_xy, uint32_t n_edges, const uint32_t *edge_start,
                   const uint32_t *edge_end, const uint32_t *edge_off, const double *edge_xy,
                   const cs_geoms *barriers, const cs_geoms *data, double max_dist, uint32_t n_nearest,
                   cs_assignment **out, uint64_t *n_out);
int cs_assignment_copy(const cs_assignment *a, uint32_t *data_index, uint32_t *node, double *dist);
void cs_assignment_free(cs_assignment *a);

/* Number of nodes/edges; maximum out/in degree. */
int cs_graph_info(const cs_graph *g, uint32_t *n_nodes, uint32_t *n_edges, uint32_t *max_out_deg,
                  uint32_t *max_in_deg);

/* Progress counter: sources finished by the last or running call on this
 * graph. Replaces NetworkStructure.progress (rust/src/graph.rs:407-415). */
uint64_t cs_graph_progress(const cs_graph *g);

/* ---------------------------------------------------------------------- */
/* Centrality                                                             */
/* ---------------------------------------------------------------------- */

/* Thresholds one kernel launch takes. Calls with more run one launch per
 * chunk of CS_MAX_THRESHOLDS, each searching to the call's largest cutoff (the
 * reference's single search); batched and served requests with more run as one
 * job per chunk in the same launches. */
#define CS_MAX_THRESHOLDS 8

typedef struct cs_params {
  uint32_t n_thresholds;     /* >= 1 */
  const uint32_t *distances; /* [T] strictly increasing */
  const float *betas;        /* [T] */
  const uint32_t *seconds;   /* [T] */
  float speed_m_s;
  float tolerance;           /* decimal (user percent / 100), clamped >= 1e-4 */
  int compute_closeness;
  int compute_betweenness;
  uint32_t n_sources;
  const uint32_t *sources;   /* [n_sources] node indices */
  const float *source_wt;    /* [n_sources] weight per source (NaN = skip) */
  const uint8_t *eligible;   /* [N] source-eligible mask */
  int track_reach;           /* fill reach_totals */
  double betw_scale;         /* post-hoc betweenness scale (1.0 = none) */
  float angular_scaling_unit;
  float farness_scaling_offset;
} cs_params;

/* Replaces NetworkStructure.centrality_shortest (rust/src/centrality.rs:1624-1874).
 * out: f64 [7][T][N] = density, farness, cycles, harmonic, beta, betweenness,
 * betweenness_beta. reach_totals: u32 [T] (may be NULL). */
int cs_centrality_shortest(cs_graph *g, const cs_params *p, double *out, uint32_t *reach_totals);

/* Device-resident form of cs_centrality_shortest, for callers that keep inputs
 * and outputs in HBM (batching, serving, benchmarks). p->distances/betas/
 * seconds are host arrays; p->sources, p->source_wt and p->eligible are device
 * arrays, as is d_cycles_wt [n_sources] (source weight / node weight). The call
 * adds into d_acc, f64 [7][T][N] laid out like `out` (zeroed by the caller), and d_reach
 * u32 [T] (may be NULL); betw_scale is not applied. stream may be NULL (the
 * null stream). Same computation as centrality_shortest
 * (rust/src/centrality.rs:1624-1874) without host copies. All node-indexed
 * buffers are in original node order; when the graph has a device order
 * (cs_graph_set_xy) the call translates them through staging it keeps with
 * the graph and returns after the stream has finished its work. */
int cs_centrality_shortest_device(cs_graph *g, const cs_params *p, const float *d_cycles_wt, double *d_acc,
                                  uint32_t *d_reach, void *stream);

/* Bytes of the static device array that cs_centrality_shortest_device_noalloc
 * works in; each device the library runs on holds one until the process ends. */
#define CS_NOALLOC_SCRATCH_BYTES (256u << 20)

/* cs_centrality_shortest_device that allocates nothing. No reference
 * counterpart: it serves callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and synchronisation as
 * cs_centrality_shortest_device; every device buffer it needs beyond them
 * (job table, control words, retry lists, node-order staging, per-wave node
 * records and scratch) lies in a static array of CS_NOALLOC_SCRATCH_BYTES on
 * the current device. The graph must already be resident there at
 * p->speed_m_s (cs_graph_upload(g, device, speed)), else CS_EINVAL. A call
 * whose fixed parts or one workgroup's waves do not fit the array is
 * CS_ENOMEM; otherwise it runs as many waves as fit (fewer than the allocating
 * entry point on large graphs). Calls on one device share the array, which
 * keeps the last call's contents: the library serialises the calls of one
 * process on a device (each returns after its stream has finished), and work
 * the caller enqueues elsewhere must not touch the array. */
int cs_centrality_shortest_device_noalloc(cs_graph *g, const cs_params *p, const float *d_cycles_wt,
                                          double *d_acc, uint32_t *d_reach, void *stream);

/* Replaces NetworkStructure.betweenness_od_shortest (rust/src/centrality.rs:2413-2540):
 * OD-weighted betweenness from a sparse origin-destination matrix given as n_pairs
 * parallel (origin, destination, trip weight) triples, as OdMatrix(origins,
 * destinations, weights) takes them (a later pair overwrites an earlier one with
 * the same origin and destination). Only p's thresholds, speed_m_s and tolerance
 * are read. Non-live origins are skipped; an index beyond the graph is CS_EINVAL.
 * out: f64 [2][T][N] = betweenness, betweenness_beta. */
int cs_betweenness_od_shortest(cs_graph *g, const cs_params *p, uint32_t n_pairs, const uint32_t *origins,
                               const uint32_t *destinations, const float *weights, double *out);

/* Replaces NetworkStructure.centrality_simplest (rust/src/centrality.rs:1876-2132).
 * out: f64 [4][T][N] = density, farness, harmonic, betweenness. */
int cs_centrality_simplest(cs_graph *g, const cs_params *p, double *out, uint32_t *reach_totals);

/* Replaces NetworkStructure.segment_centrality (rust/src/centrality.rs:2134-2407).
 * out: f64 [4][T][N] = segment density, harmonic, beta, betweenness. */
int cs_segment_centrality(cs_graph *g, const cs_params *p, double *out);

/* ---------------------------------------------------------------------- */
/* Data layer                                                             */
/* ---------------------------------------------------------------------- */

/* Replaces DataMap.accessibility (rust/src/data.rs:617-747) over the
 * assignment DataMap.node_data_map holds after
 * assign_data_to_network (data.rs:381-387, 455-512), given here as n_assign
 * (node, entry, distance) triples. Entry e (0..n_entries-1) belongs to dedupe
 * group entry_group[e] (< n_entries; entries without a dedupe key form groups
 * of their own) and land-use class entry_class[e] (0..n_classes-1, or -1 when
 * its class is not among the accessibility keys). Per live source node the
 * nearest member of each group within max_seconds * speed_m_s counts
 * (data.rs:520-615); equal totals within a group go to the member given first.
 * distances / betas / max_curve_wts: n_thresholds values as
 * pair_distances_betas_time and clip_wts_curve return them; max_seconds is
 * the largest of the thresholds' seconds. angular != 0 searches simplest-path
 * trees on a dual graph instead of shortest-path trees (data.rs:538-548;
 * CS_EINVAL on a primal graph, as validate_dual_for_angular raises).
 * out_count, out_weighted: f64 [n_classes][n_thresholds][N] (unweighted and
 * weighted accessibility); out_nearest: f64 [n_classes][N], the distance to
 * the nearest counted entry within the largest distance (NaN: none). */
int cs_accessibility(cs_graph *g, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                     uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node,
                     const uint32_t *assign_entry, const double *assign_dist, uint32_t n_thresholds,
                     const uint32_t *distances, const float *betas, const float *max_curve_wts,
                     uint32_t max_seconds, float speed_m_s, int angular, double *out_count,
                     double *out_weighted, double *out_nearest);

/* Replaces DataMap.mixed_uses (rust/src/data.rs:749-944): the same inputs as
 * cs_accessibility (angular included), with entry_class numbering every land
 * use in sorted order (n_classes = the number of distinct land uses). out:
 * f64 [8][n_thresholds][N] = hill q0, q1, q2, distance-weighted hill q0, q1,
 * q2, shannon, gini-simpson (diversity.rs:6-175, 341-385), each computed in
 * f32 from the per-source class counts and nearest class distances. */
int cs_mixed_uses(cs_graph *g, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                  uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node, const uint32_t *assign_entry,
                  const double *assign_dist, uint32_t n_thresholds, const uint32_t *distances, const float *betas,
                  const float *max_curve_wts, uint32_t max_seconds, float speed_m_s, int angular, double *out);

/* Replaces DataMap.stats (rust/src/data.rs:1013-1233): the entries, dedupe
 * groups, node assignment and angular flag of cs_accessibility, and
 * n_maps numeric maps as values f32 [n_maps][n_entries] (NaN: the entry has no
 * value). out: f64 [n_maps][14][n_thresholds][N] = sum, sum_wt, mean,
 * mean_wt, median, median_wt, count, count_wt, variance, variance_wt, mad,
 * mad_wt, max, min per live source, over the entries it reaches within each
 * distance (nodes that are not live keep 0 for sums and counts, NaN for the
 * rest, as StatsResult::new starts them). Sums run in f32 in ascending entry
 * order (the reference's order is its hash map's, which varies from run to
 * run); medians follow the reference's total order and stable sort. */
int cs_data_stats(cs_graph *g, uint32_t n_entries, const uint32_t *entry_group, uint32_t n_assign,
                  const uint32_t *assign_node, const uint32_t *assign_entry, const double *assign_dist,
                  uint32_t n_maps, const float *values, uint32_t n_thresholds, const uint32_t *distances,
                  const float *betas, const float *max_curve_wts, uint32_t max_seconds, float speed_m_s,
                  int angular, double *out);

/* ---------------------------------------------------------------------- */
/* Single-source trees                                                    */
/* ---------------------------------------------------------------------- */

/* Replaces NetworkStructure.dijkstra_tree_shortest (rust/src/centrality.rs:1499-1508).
 * Arrays of length N; pred = -1 for none; order receives the settle order. */
int cs_tree_shortest(cs_graph *g, uint32_t src, uint32_t max_seconds, float speed_m_s,
                     uint8_t *visited, uint8_t *discovered, int32_t *pred, float *short_dist,
                     float *agg_seconds, uint32_t *order, uint32_t *n_order);

/* Replaces NetworkStructure.dijkstra_tree_simplest (rust/src/centrality.rs:1510-1521). */
int cs_tree_simplest(cs_graph *g, uint32_t src, uint32_t max_seconds, float speed_m_s,
                     uint8_t *visited, uint8_t *discovered, int32_t *pred, float *simpl_dist,
                     float *agg_seconds, uint32_t *order, uint32_t *n_order);

/* Replaces NetworkStructure.dijkstra_tree_segment (rust/src/centrality.rs:1523-1611).
 * origin_seg / last_seg hold edge ids (-1 = none); edge_order receives the
 * visited edges in visit order. */
int cs_tree_segment(cs_graph *g, uint32_t src, uint32_t max_seconds, float speed_m_s,
                    uint8_t *visited, uint8_t *discovered, int32_t *pred, float *short_dist,
                    float *agg_seconds, int32_t *origin_seg, int32_t *last_seg, uint32_t *order,
                    uint32_t *n_order, uint32_t *edge_order, uint32_t *n_edges);

/* Multi-predecessor Brandes state of one shortest-path source, the state the
 * reference builds in dijkstra_brandes_shortest (rust/src/centrality.rs:1344-1494):
 * secs/route f32 [N], sigma f64 [N], n_preds u32 [N], preds u32 [N*max_preds]
 * (predecessor node ids in insertion order). Exposed for parity checks. */
int cs_brandes_shortest(cs_graph *g, uint32_t src, uint32_t max_seconds, float speed_m_s,
                        float tolerance, float *secs, float *route, double *sigma,
                        uint32_t *n_preds, uint32_t *preds, uint32_t max_preds);

/* ---------------------------------------------------------------------- */
/* Measurement                                                            */
/* ---------------------------------------------------------------------- */

typedef struct cs_stats {
  double kernel_ms;        /* HIP-event time of the last call's search kernel(s) */
  double total_ms;         /* host wall time of the last call */
  uint64_t sources;        /* sources searched */
  uint64_t nodes_reached;  /* sum over sources of R_s */
  uint64_t edges_scanned;  /* sum over sources of traversed in-edges E_s */
  uint64_t within_sum;     /* sum over sources and thresholds of R_{s,i} */
  uint32_t waves;          /* persistent waves launched */
  uint32_t retries;        /* sources re-run with larger workspaces */
} cs_stats;

/* Statistics of the calling thread's last compute call. */
int cs_last_stats(cs_stats *out);

/* ---------------------------------------------------------------------- */
/* Launch shape and batching                                              */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference, which runs one rayon loop per call
 * (rust/src/centrality.rs:1702-1836); these entry points serve callers that
 * run many centrality requests (many graphs, many threshold sets) at once. */

typedef struct cs_tuning {
  uint32_t block_threads;   /* workgroup size 64..256, a multiple of 64: sources in flight per workgroup = /64 */
  uint32_t waves_per_simd;  /* cap on resident waves per SIMD; 0 = as many as fit */
  float delta_scale;        /* bucket width as a fraction in (0, 1] of its safe maximum */
  uint32_t slot_capacity;   /* settled-node slots per wave (simplest: search states per lane) before a
                               retry with four times as many; 0 = default (16384; simplest 8192). For
                               simplest paths 1/128 of it (2..64, default 64) also bounds the entries of
                               one bucket of a lane's route-cost queue, beyond which the source runs
                               again with a heap */
  uint32_t xcd_split;       /* 1: each XCD's waves take a contiguous share of the sources first; 0: one counter */
  uint32_t simplest_lanes;  /* simplest: sources per wave (1..64), 0 = chosen from the launch's source count */
} cs_tuning;

/* Launch shape of the plain shortest/segment entry points (process-wide). */
int cs_set_tuning(const cs_tuning *t);
int cs_get_tuning(cs_tuning *t);

/* No counterpart in the reference (its searches allocate per call). Frees the
 * search workspaces the library pools per device between calls (node records,
 * per-wave scratch, launch tables), waiting for calls that hold them; the next
 * call allocates fresh, zeroed ones. For callers that want the HBM back. */
int cs_release_workspaces(void);

/* A batch of node_centrality_shortest jobs (any graphs, each with its own
 * thresholds and sources) run as one persistent launch captured in a hipGraph.
 * Graphs must outlive the batch. Not thread-safe: one thread drives a batch. */
typedef struct cs_batch cs_batch;
int cs_batch_create(cs_batch **out);
void cs_batch_destroy(cs_batch *b);
/* Adds a job with the arguments of cs_centrality_shortest, any number of
 * thresholds; `out` ([7][T][N]) and `reach_totals` ([T], may be NULL) are
 * filled by cs_batch_results. Returns the job index (>= 0, in call order) or
 * minus a status code. */
int cs_batch_add_shortest(cs_batch *b, cs_graph *g, const cs_params *p, double *out, uint32_t *reach_totals);
/* Times candidate launch shapes on a `sample_fraction` of the batch's sources
 * and keeps the fastest; `chosen` (may be NULL) receives it. */
int cs_batch_autotune(cs_batch *b, double sample_fraction, cs_tuning *chosen);
int cs_batch_set_tuning(cs_batch *b, const cs_tuning *t);
/* Runs every job: the first run uploads the jobs and captures the launch,
 * later runs replay it. Results stay in device memory until cs_batch_results. */
int cs_batch_run(cs_batch *b);
/* Copies the last run's results into the jobs' `out` / `reach_totals` arrays. */
int cs_batch_results(cs_batch *b);

/* ---------------------------------------------------------------------- */
/* Serving: a request loop with continuous batching                       */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference. Requests are admitted first-come
 * first-served under an HBM budget; admitted shortest-path requests share
 * back-to-back batched launches (each launch takes part of every active
 * request's remaining sources, so requests join and leave between launches);
 * simplest-path requests run on a second stream. Graphs must outlive the
 * requests that use them; `out` / `reach_totals` must stay valid until the
 * request completes. */

#define CS_REQ_SHORTEST 0 /* out [7][T][N], as cs_centrality_shortest */
#define CS_REQ_SIMPLEST 1 /* out [4][T][N], as cs_centrality_simplest */

typedef struct cs_server cs_server;
typedef struct cs_server_config {
  uint64_t hbm_budget_bytes; /* device bytes admitted requests may hold; 0 = a quarter of free HBM */
  uint32_t max_active;       /* requests on the device at once; 0 = 64 */
  double slice_ms;           /* target duration of one batched launch; 0 = 100 ms */
} cs_server_config;
typedef struct cs_request_info {
  int32_t status;     /* 0 queued, 1 running, 2 done, or minus the error status */
  double queued_ms;   /* submit -> admission (so far, if still queued) */
  double latency_ms;  /* submit -> results in `out` (so far, if not done) */
  uint64_t sources;   /* sources to search */
  uint32_t launches;  /* batched launches the request took part in */
} cs_request_info;

int cs_server_create(const cs_server_config *cfg, cs_server **out);
/* Stops after the running launches; queued requests fail with CS_EBUSY. No
 * thread may be inside cs_server_wait on this server. */
void cs_server_destroy(cs_server *s);
/* Returns a request id (>= 0) or minus a status code. */
int64_t cs_server_submit(cs_server *s, cs_graph *g, int kind, const cs_params *p, double *out,
                         uint32_t *reach_totals);
int cs_server_poll(cs_server *s, int64_t id, cs_request_info *info);
/* Blocks until the request completes; returns its status. A request whose
 * worker failed (a HIP error, an exception) completes with CS_EINTERNAL (or
 * CS_ENOMEM for host memory); later submissions of that kind then fail at
 * once with the same message. */
int cs_server_wait(cs_server *s, int64_t id, cs_request_info *info);
/* cs_server_wait for at most timeout_ms (negative: no limit); CS_ETIMEDOUT if
 * the request is still queued or running then (info holds its state). */
int cs_server_wait_for(cs_server *s, int64_t id, double timeout_ms, cs_request_info *info);
/* Cancels a request: a queued one completes at once, a running one once its
 * sources already in a launch have finished (no more are handed out); either
 * way with status CS_ECANCELED and its device memory released. A request that
 * has already completed is left as it is. */
int cs_server_cancel(cs_server *s, int64_t id);
/* Forgets a completed request. */
int cs_server_release(cs_server *s, int64_t id);
/* Failure-handling tests: makes the worker of `kind` fail its next slice with a
 * std::runtime_error (what = 1) or std::bad_alloc (2), or end itself with an
 * exception outside the per-request handling (3). No reference counterpart. */
int cs_server_inject_fault(cs_server *s, int kind, int what);

#ifdef __cplusplus
}
#endif
#endif
