// cs_internal.hpp -- host-side graph representation and kernel launch
// interfaces shared by the host library (cs_host.cpp) and the HIP kernels
// (cs_search.hip, cs_angular.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cityseer_amd.h"

namespace cs {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 independent waves per workgroup
constexpr uint32_t kMaxT = CS_MAX_THRESHOLDS;
constexpr uint32_t kNodeMask = 0x7fffffffu;
constexpr uint32_t kCanonBit = 0x80000000u;
constexpr float kTieEps = 1e-4f;  // centrality.rs:28
constexpr float kAngularTieBreak = 1e-6f;  // centrality.rs:24

// Thrown by host code; converted to a status + thread-local message at the C ABI.
struct Error {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg);
void hip_check(hipError_t e, const char *what);

// Device memory of n elements owned by one scope (freed on exit, errors included).
template <typename T>
struct DevMem {
  T *p = nullptr;
  explicit DevMem(size_t n) { hip_check(hipMalloc(&p, sizeof(T) * (n ? n : 1)), "hipMalloc"); }
  ~DevMem() { (void)hipFree(p); }
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
};

// Speed-dependent device edge arrays for one (device, speed).
struct DevEdges {
  float speed = 0.0f;
  // shortest / segment (reverse search over in-edges, fold over out-edges)
  // bit 31 marks the entry that counts its undirected key toward the circuit
  // rank: in-edges whose key's canonical edge is this edge or its reverse twin,
  // out-edges that are canonical one-way edges
  uint2 *rev_adj = nullptr;     // {nb | count<<31, bits(seconds of nb->v)}
  uint2 *fwd_adj = nullptr;     // {u  | count<<31, bits(seconds of v->u)}
  float2 *rev_tw = nullptr;     // {twin length, twin imp} per rev position (segment)
  float *fwd_twlen = nullptr;   // twin length per fwd position (segment)
  uint32_t *rev_eid = nullptr;  // edge id per rev position (trees)
  uint32_t *fwd_eid = nullptr;  // edge id per fwd position (trees)
  bool has_out_canon = false;   // some out-edge carries the circuit-rank bit
  float min_es = 0.0f;          // min seconds over non-loop edges (use_impedance)
  float max_es = 0.0f;
  // angular (forward search over out-edges on the 2-state dual)
  uint4 *ang_adj = nullptr;  // {v | srcslot<<30 | dstslot<<31 ... see cs_angular.hip}
  uint4 *ang_srow = nullptr;  // [2N][4] usable out-edges of each dual state (cs_angular.hip)
  float min_es_ang = 0.0f;
  float max_w_ang = 0.0f;  // largest route increment (angle + 1e-6 length) of the dual's edges
};

struct DevGraph {
  int device = -1;
  uint32_t *rev_off = nullptr;  // original node order (angular searches on dual graphs)
  uint32_t *fwd_off = nullptr;
  uint4 *rows = nullptr;  // {fwd begin, fwd end, rev begin, rev end} per device node: one 16-byte load
  uint32_t *orig = nullptr;  // [N] original index of each device node (tie order)
  uint32_t *perm = nullptr;  // [N] device index of each original node
  // staging for calls whose device buffers are in original order (grown on
  // demand, used under the graph's lock)
  void *xbuf = nullptr;
  size_t xbuf_bytes = 0;
  std::vector<std::unique_ptr<DevEdges>> edges;  // small cache keyed by speed
};

struct Graph {
  uint32_t n = 0, m = 0;
  bool is_dual = false;
  std::vector<double> z;
  std::vector<uint8_t> live;
  std::vector<float> weight;
  std::vector<uint32_t> esrc, edst, ekey;
  std::vector<float> elen, eimp, eang, esec;
  std::vector<int32_t> epk;
  // topology in petgraph adjacency order (newest edge first)
  std::vector<uint32_t> rev_off, rev_eid;  // in-edges per node
  std::vector<uint32_t> fwd_off, fwd_eid;  // out-edges per node
  std::vector<int32_t> twin;                // edges_connecting(dst, src).find(key)
  std::vector<uint8_t> canon;               // first edge of its (min,max,key) triple
  std::vector<int8_t> slot_src, slot_dst;   // dual endpoint slots (centrality.rs:533-566)
  bool slots_ok = false;
  std::string slots_err;
  uint32_t max_out_deg = 0, max_in_deg = 0;
  bool has_loops = false;
  // Device node order: nodes lie in device memory along a space-filling curve
  // of their coordinates when those are given (primal graphs), so that nodes
  // near in space are near in memory. perm[original] = device index, iperm =
  // its inverse; both identity otherwise. Every entry point translates.
  std::vector<uint32_t> perm, iperm;
  bool identity_order = true;
  mutable std::atomic<bool> order_frozen{false};  // set once an entry point has translated ids

  std::mutex mu;
  std::map<int, std::unique_ptr<DevGraph>> dev;
  // Live progress (sources finished by the running or last call): sources
  // the host settles itself plus one count per wave (or lane) that the kernels
  // keep in pinned host memory while they run (progress_begin).
  std::atomic<uint32_t *> progress_slots{nullptr};  // [kProgressSlots], host address
  uint32_t *progress_dev = nullptr;                 // the same memory, device address
  std::atomic<uint64_t> progress_host{0};
};

// Counts the kernels keep for a call's live progress: one per wave (search)
// or per lane (angular search); launches wider than this report none.
constexpr uint32_t kProgressSlots = 65536;

// Starts a call's live progress on g (under g.mu, no kernel of g running):
// clears the counts, counts `settled_on_host` sources at once and returns the
// device address of the counts for the launches to pass on.
uint32_t *progress_begin(Graph &g, uint64_t settled_on_host);
// Ends a call's progress at n sources: whatever the kernels have not counted
// is added on the host (calls that report only on completion count it all).
void progress_finish(Graph &g, uint64_t n);

// Frees the pooled search workspaces of every device, waiting for the calls
// that hold them; the next call allocates and zeroes fresh ones.
void release_search_workspaces();
void release_angular_workspaces();

// Host edge travel seconds (centrality.rs:970-1007).
float edge_seconds(const Graph &g, uint32_t from, uint32_t to, uint32_t e, float speed,
                   bool use_impedance);

// Device node order from coordinates (cs_host.cpp); before any device use.
void set_node_order_from_xy(Graph &g, const double *x, const double *y);

// Device residency (cs_host.cpp).
DevGraph &device_graph(Graph &g, int device);
DevEdges &device_edges(Graph &g, DevGraph &dg, float speed);
// The same, without staging anything: null when not resident yet.
DevGraph *find_device_graph(Graph &g, int device);
DevEdges *find_device_edges(DevGraph &dg, float speed);

// Per-call statistics (thread-local).
cs_stats &thread_stats();

// ----------------------------------------------------------------------
// Kernel launchers (cs_search.hip)
// ----------------------------------------------------------------------

struct CallSpec {
  uint32_t T = 0;
  float dist[kMaxT] = {0};
  float beta[kMaxT] = {0};
  uint32_t seconds[kMaxT] = {0};
  float maxsec = 0.0f;
  float speed = 0.0f;
  float tolerance = 1e-4f;
  bool closeness = true, betweenness = true, track_reach = false;
  float ang_unit = 180.0f, far_offset = 1.0f;
};

// Shortest-path centrality over `n_sources` sources; acc is a device array
// [7][T][N] (metric-major, the reference's row order) that the call
// accumulates into; reach [T] u32.
// tstride: threshold stride of d_acc's rows (0: cs.T); a threshold chunk
// passes its first row and the call's threshold count.
// progress: the call's live counts (progress_begin), or null.
void launch_shortest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                     const uint32_t *d_sources, const float *d_wt, const float *d_cyc,
                     uint32_t n_sources, const uint8_t *d_eligible, double *d_acc,
                     uint32_t *d_reach, hipStream_t stream, cs_stats &st, uint32_t tstride = 0,
                     uint32_t *progress = nullptr);

// launch_shortest with sources, eligibility and acc in original node order
// (device buffers of the caller), translated to and from the graph's device
// order through staging kept with the device graph (dg.xbuf); the caller
// holds the graph's lock. Returns after the stream has finished the call's
// work when the order is not the identity.
void launch_shortest_original(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                              const uint32_t *d_sources, const float *d_wt, const float *d_cyc,
                              uint32_t n_sources, const uint8_t *d_eligible, double *d_acc,
                              uint32_t *d_reach, hipStream_t stream, cs_stats &st, uint32_t tstride = 0,
                              uint32_t *progress = nullptr);

// launch_shortest_original with every device buffer beyond the caller's
// inputs and outputs carved from the device region ws (ws_bytes bytes,
// 256-byte aligned): no allocation. The call's rows go to thresholds
// t0 .. t0 + cs.T - 1 of d_acc, f64 [7][tstride][N]. Returns after the stream
// has finished the call's work.
void launch_shortest_in(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs, const uint32_t *d_sources,
                        const float *d_wt, const float *d_cyc, uint32_t n_sources, const uint8_t *d_eligible,
                        double *d_acc, uint32_t *d_reach, uint32_t tstride, uint32_t t0, hipStream_t stream,
                        cs_stats &st, char *ws, size_t ws_bytes);

// Launch shape of the search kernel: workgroup size (64..256), a cap on
// resident waves per SIMD (0 = as many as fit), the bucket width as a fraction
// of its safe maximum (<= 1) and the per-wave slot capacity (0 = default).
struct Tuning {
  uint32_t block = 256;
  uint32_t waves_per_simd = 0;
  float delta_scale = 1.0f;
  uint32_t cap = 0;
  uint32_t xcd_split = 1;  // sources dealt to XCDs in contiguous shares (1) or from one counter (0)
  uint32_t ang_lanes = 0;  // angular: sources per wave (1..64), 0 = from the launch's source count
};
void set_default_tuning(const Tuning &t);
Tuning default_tuning();

// One (graph, request) pair of a batched shortest-path launch. d_acc is
// [7][tstride][N] with tstride >= cs.T (0: the launch's NT); thresholds beyond
// cs.T stay zero.
struct JobSpec {
  Graph *g;
  DevGraph *dg;
  DevEdges *de;
  CallSpec cs;
  const uint32_t *d_sources;
  const float *d_wt;
  const float *d_cyc;
  uint32_t n_sources;
  const uint8_t *d_eligible;
  double *d_acc;
  uint32_t *d_reach;
  float delta_scale;
  uint32_t tstride = 0;
  // OD-weighted betweenness: source position i seeds od_w at od_dest over
  // [od_off[i], od_off[i + 1]) instead of every reached node (device arrays)
  const uint32_t *d_od_off = nullptr;
  const uint32_t *d_od_dest = nullptr;
  const float *d_od_w = nullptr;
  // per-wave records beyond the graph's nodes (land-use dedupe group tables)
  uint32_t extra_records = 0;
};
// All sources of all jobs in one persistent launch, or only `items`
// ({job, source position}) when given.
void launch_shortest_batch(const std::vector<JobSpec> &jobs, uint32_t NT, const Tuning &tune,
                           hipStream_t stream, cs_stats &st, const std::vector<uint2> *items = nullptr,
                           uint32_t *progress = nullptr);

// A prepared batch of shortest-path jobs replayed as one captured hipGraph
// (cs_search.hip). The zero regions (64-bit words) are cleared inside the
// graph before each search: the jobs' accumulators and reach counters.
struct ZeroRegion {
  uint64_t *ptr;
  size_t words;
};
struct ShortestBatch;
ShortestBatch *batch_new(std::vector<JobSpec> specs, std::vector<ZeroRegion> zero, uint32_t NT, int device,
                         const Tuning &tune);
void batch_delete(ShortestBatch *b);
void batch_set_tuning(ShortestBatch &b, const Tuning &t);
Tuning batch_tuning(const ShortestBatch &b);
Tuning batch_autotune(ShortestBatch &b, double fraction, hipStream_t stream);
void batch_run(ShortestBatch &b, hipStream_t stream, cs_stats &st);

// Segment centrality; acc [4][T][N].
void launch_segment(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                    const uint32_t *d_sources, uint32_t n_sources, double *d_acc,
                    hipStream_t stream, cs_stats &st, uint32_t tstride = 0, uint32_t *progress = nullptr);

// Single-source outputs (trees and Brandes state), device arrays of length N.
struct TreeOut {
  float *secs = nullptr;       // agg seconds (INF = unreached)
  int32_t *pred = nullptr;     // pred node (-1)
  int32_t *last_eid = nullptr; // edge id of v->pred (-1)
  int32_t *origin_eid = nullptr;
  double *sigma = nullptr;     // Brandes sigma
  uint32_t *pred_list = nullptr;  // Brandes preds [N * max_preds]
  uint32_t *n_preds = nullptr;
  uint32_t max_preds = 0;
};
// Land-use accessibility over each source's shortest-path tree (device
// arrays, nodes in device order). A member is one assignment of a data entry
// to a node; the members of one dedupe group lie together.
struct AccessOut {
  const uint32_t *node_off = nullptr;    // [N + 1] into node_mem
  const uint32_t *node_mem = nullptr;    // member indices assigned to each node
  const uint4 *mem = nullptr;            // {node, entry, data distance f32 bits, shared group slot or kSoloGroup}
  uint32_t n_shared = 0;                 // dedupe groups of two or more members (one table slot each)
  const int32_t *entry_class = nullptr;  // class of each entry (-1: not counted)
  double *count = nullptr;               // [C][T][N] entries within each distance
  double *weighted = nullptr;            // [C][T][N] clipped negative-exponential weights
  unsigned long long *nearest = nullptr; // [C][N] f64 bits of the nearest entry (NaN bits: none)
  // land-use statistics instead of counts (data.rs:1013-1233): when stats is
  // set, each source's rows [map][kStatMetrics][t] are computed from the
  // numeric values of the entries it reaches, [maps][n_entries] (NaN: none)
  double *stats = nullptr;
  const float *values = nullptr;
  uint32_t n_maps = 0, n_entries = 0;
  uint32_t n_groups = 0;  // dedupe groups with at least one member (winners per source at most)
  uint32_t T = 0, N = 0;
  uint32_t tstride = 0;   // threshold stride of the count / weighted rows (and of the diversity rows)
  uint32_t max_mask = 0;  // thresholds equal to the largest distance
  float maxdist = 0.0f;   // max walk seconds * speed
  float dist[kMaxT], beta[kMaxT], mcw[kMaxT];
};
// A member whose dedupe group has no other member.
constexpr uint32_t kSoloGroup = 0xffffffffu;
void launch_access(Graph &g, DevGraph &dg, DevEdges &de, float maxsec, const uint32_t *d_sources,
                   uint32_t n_sources, const AccessOut &out, hipStream_t stream, cs_stats &st,
                   uint32_t *progress = nullptr);
// Diversity of each source's class counts (the rows launch_access leaves in
// out.count / out.nearest, n_classes of them): div [kDivMetrics][T][N] =
// hill q0, q1, q2, distance-weighted hill q0, q1, q2, shannon, gini-simpson.
constexpr uint32_t kDivMetrics = 8;
// Statistics rows per numeric map, in this order: sum, sum_wt, mean, mean_wt,
// median, median_wt, count, count_wt, variance, variance_wt, mad, mad_wt, max, min.
constexpr uint32_t kStatMetrics = 14;
void launch_diversity(const AccessOut &out, uint32_t n_classes, double *div, hipStream_t stream);

// mode 0: single-predecessor tree, 1: Brandes multi-predecessor state.
void launch_single(Graph &g, DevGraph &dg, DevEdges &de, uint32_t src, float maxsec,
                   float tolerance, int mode, const TreeOut &out, hipStream_t stream);

// Angular (simplest) centrality (cs_angular.hip); acc [N][4][T].
// One (graph, request) pair of an angular launch; sources index d_sources.
struct AngSpec {
  Graph *g;
  DevGraph *dg;
  DevEdges *de;
  CallSpec cs;
  const uint32_t *d_sources;
  const float *d_wt;
  const uint8_t *d_eligible;
  double *d_acc;  // [4][tstride][N]
  uint32_t *d_reach;
  uint32_t tstride = 0;  // threshold stride of d_acc (0: cs.T)
};

// Simplest-path centrality for several requests in one launch: the items
// {job, source position} (or, without a list, sources 0..n_implicit-1 of job 0).
void launch_simplest_jobs(const std::vector<AngSpec> &specs, const std::vector<uint2> *explicit_items,
                          uint32_t n_implicit, hipStream_t stream, cs_stats &st, uint32_t *progress = nullptr);

void launch_simplest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                     const uint32_t *d_sources, const float *d_wt, uint32_t n_sources,
                     const uint8_t *d_eligible, double *d_acc, uint32_t *d_reach,
                     hipStream_t stream, cs_stats &st, uint32_t tstride = 0, uint32_t *progress = nullptr);

// dijkstra_tree_simplest for one source (cs_angular.hip); device arrays of
// length N, pre-filled by the caller (INF / -1 / 0).
void launch_tree_simplest(Graph &g, DevGraph &dg, DevEdges &de, uint32_t src, float maxsec, float *d_simpl,
                          float *d_secs, int32_t *d_pred, uint8_t *d_vis, uint32_t *d_order, uint32_t *d_norder,
                          hipStream_t stream);

// Reached-node lists of many simplest-path trees (cs_angular.hip), for angular
// data aggregation: item i searches from d_src[d_items[i]] (d_src[i] without
// d_items) and leaves {node | seconds << 32} of the nodes it reaches, in
// first-reach order, in d_lists[i * cap ...] and their number in d_counts[i].
// Items whose search ran out of room (more than cap nodes or 2 cap states, or a
// full queue: an overfilled bucket of the bucket ring, or more than hcap
// entries of the heap that `heap` selects) get kNoList; their positions are
// returned, the causes (kReachFailCap, kReachFailHeap) in *fail_bits.
constexpr uint32_t kNoList = 0xffffffffu;
constexpr uint32_t kReachFailCap = 1u, kReachFailHeap = 2u;
std::vector<uint32_t> launch_reach_simplest(Graph &g, DevGraph &dg, DevEdges &de, float maxsec,
                                            const uint32_t *d_src, const uint32_t *d_items, uint32_t n_items,
                                            uint32_t cap, uint32_t hcap, bool heap, uint64_t *d_lists,
                                            uint32_t *d_counts, hipStream_t stream, cs_stats &st,
                                            uint32_t *fail_bits);

// Data aggregation (accessibility rows, class counts or statistics, as `out`
// asks) over reached-node lists (cs_search.hip): item i is source
// d_src[d_items[i]] with the list d_lists[i * lcap ...] of d_counts[i] entries;
// items whose count is kNoList are skipped.
void launch_access_lists(const AccessOut &out, const uint32_t *d_src, const uint32_t *d_items,
                         const uint64_t *d_lists, const uint32_t *d_counts, uint32_t n_items, uint32_t lcap,
                         float maxsec, float speed, hipStream_t stream, cs_stats &st, uint32_t *progress);

}  // namespace cs
