// cs_host.cpp -- host side of the C ABI: graph construction, device
// residency, threshold conversion and the extern "C" entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <numeric>
#include <unordered_map>

#include "cs_api.hpp"

namespace cs {

namespace {
thread_local std::string t_err;
thread_local cs_stats t_stats;
}  // namespace

void fail(int code, const std::string &msg) { throw Error{code, msg}; }

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? CS_ENOMEM : CS_EDEVICE;
    (void)hipGetLastError();
    fail(code, std::string(what) + ": " + hipGetErrorString(e));
  }
}

cs_stats &thread_stats() { return t_stats; }

// Tobler slope penalty and edge travel seconds (centrality.rs:970-1007).
float edge_seconds(const Graph &g, uint32_t from, uint32_t to, uint32_t e, float speed,
                   bool use_impedance) {
  if (!std::isnan(g.esec[e])) return g.esec[e];
  const float length = g.elen[e];
  float slope_pen = 1.0f;
  if (length > 0.0f) {
    const double zf = g.z[from], zt = g.z[to];
    if (!std::isnan(zf) && !std::isnan(zt)) {
      const float slope = (float)(zt - zf) / length;
      const float flat = 0.839457f;
      const float factor = std::exp(-3.5f * std::fabs(slope + 0.05f));
      slope_pen = flat / factor;
    }
  }
  const float imp = use_impedance ? g.eimp[e] : 1.0f;
  float s = length * imp;
  s = s * slope_pen;
  return s / speed;
}

namespace {

// Build the petgraph-ordered adjacency (newest edge first), twins, canonical
// flags and the dual endpoint slots.
void build_topology(Graph &g) {
  const uint32_t n = g.n, m = g.m;
  g.rev_off.assign(n + 1, 0);
  g.fwd_off.assign(n + 1, 0);
  for (uint32_t e = 0; e < m; ++e) {
    g.fwd_off[g.esrc[e] + 1]++;
    g.rev_off[g.edst[e] + 1]++;
  }
  for (uint32_t i = 0; i < n; ++i) {
    g.max_out_deg = std::max(g.max_out_deg, g.fwd_off[i + 1]);
    g.max_in_deg = std::max(g.max_in_deg, g.rev_off[i + 1]);
    g.fwd_off[i + 1] += g.fwd_off[i];
    g.rev_off[i + 1] += g.rev_off[i];
  }
  g.fwd_eid.resize(m);
  g.rev_eid.resize(m);
  std::vector<uint32_t> fpos(g.fwd_off.begin(), g.fwd_off.end() - 1);
  std::vector<uint32_t> rpos(g.rev_off.begin(), g.rev_off.end() - 1);
  // descending edge id within each row: petgraph lists the newest edge first
  for (uint32_t k = m; k-- > 0;) {
    g.fwd_eid[fpos[g.esrc[k]]++] = k;
    g.rev_eid[rpos[g.edst[k]]++] = k;
  }
  g.has_loops = false;
  for (uint32_t e = 0; e < m; ++e)
    if (g.esrc[e] == g.edst[e]) g.has_loops = true;
  // twin(e) = edges_connecting(dst, src).find(key) (graph.rs:1278-1292)
  g.twin.assign(m, -1);
  for (uint32_t e = 0; e < m; ++e) {
    const uint32_t a = g.edst[e], b = g.esrc[e], key = g.ekey[e];
    for (uint32_t j = g.fwd_off[a]; j < g.fwd_off[a + 1]; ++j) {
      const uint32_t f = g.fwd_eid[j];
      if (g.edst[f] == b && g.ekey[f] == key) {
        g.twin[e] = (int32_t)f;
        break;
      }
    }
  }
  // canonical edge of each undirected key (min, max, edge_idx); loops excluded
  // (circuit_ranks_from_traversal de-duplicates with a HashSet, centrality.rs:486-506)
  g.canon.assign(m, 0);
  {
    std::unordered_map<uint64_t, std::vector<std::pair<uint32_t, uint32_t>>> seen;
    seen.reserve(m);
    for (uint32_t e = 0; e < m; ++e) {
      const uint32_t a = g.esrc[e], b = g.edst[e];
      if (a == b) continue;
      const uint32_t lo = std::min(a, b), hi = std::max(a, b);
      const uint64_t hk = ((uint64_t)lo << 32) | hi;
      auto &lst = seen[hk];
      bool dup = false;
      for (auto &kv : lst)
        if (kv.first == g.ekey[e]) dup = true;
      if (!dup) {
        lst.emplace_back(g.ekey[e], e);
        g.canon[e] = 1;
      }
    }
  }
  // dual endpoint slots over edges in index order (centrality.rs:533-566)
  g.slot_src.assign(m, -1);
  g.slot_dst.assign(m, -1);
  g.slots_ok = true;
  std::vector<int32_t> slots(2 * (size_t)n, -1);
  for (uint32_t e = 0; e < m && g.slots_ok; ++e) {
    const int32_t k = g.epk[e];
    if (k < 0) {
      g.slots_ok = false;
      g.slots_err = "dual edge " + std::to_string(g.esrc[e]) + " -> " + std::to_string(g.edst[e]) +
                    " is missing shared_primal_node_key metadata";
      break;
    }
    for (uint32_t node : {g.esrc[e], g.edst[e]}) {
      int32_t *s = &slots[2 * (size_t)node];
      if (s[0] == k || s[1] == k) continue;
      if (s[0] < 0)
        s[0] = k;
      else if (s[1] < 0)
        s[1] = k;
      else {
        g.slots_ok = false;
        g.slots_err = "dual node " + std::to_string(node) + " references more than two primal endpoints";
        break;
      }
    }
  }
  if (g.slots_ok) {
    for (uint32_t e = 0; e < m; ++e) {
      const int32_t k = g.epk[e];
      const int32_t *ss = &slots[2 * (size_t)g.esrc[e]];
      const int32_t *sd = &slots[2 * (size_t)g.edst[e]];
      g.slot_src[e] = (int8_t)(ss[0] == k ? 0 : 1);
      g.slot_dst[e] = (int8_t)(sd[0] == k ? 0 : 1);
    }
  }
}

template <typename T>
T *dev_upload(const std::vector<T> &v) {
  T *p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * std::max<size_t>(v.size(), 1)), "hipMalloc graph");
  if (!v.empty()) hip_check(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice), "upload");
  return p;
}

void free_edges(DevEdges &d) {
  (void)hipFree(d.rev_adj);
  (void)hipFree(d.fwd_adj);
  (void)hipFree(d.rev_tw);
  (void)hipFree(d.fwd_twlen);
  (void)hipFree(d.rev_eid);
  (void)hipFree(d.fwd_eid);
  (void)hipFree(d.ang_adj);
  (void)hipFree(d.ang_srow);
}

}  // namespace

// Hilbert curve index of (x, y) on a 2^16 x 2^16 grid (the classic xy-to-d
// walk: quadrant by quadrant, rotating the frame as the curve does).
static uint64_t hilbert_index(uint32_t x, uint32_t y) {
  const uint32_t n = 1u << 16;
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s >>= 1) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

void set_node_order_from_xy(Graph &g, const double *x, const double *y) {
  if (!g.dev.empty() || g.order_frozen)
    fail(CS_EINVAL, "node coordinates must be set before the graph's first computation");
  const uint32_t n = g.n;
  if (g.is_dual || n < 2) return;  // dual graphs keep their order (angular search)
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (uint32_t i = 0; i < n; ++i) {
    if (!std::isfinite(x[i]) || !std::isfinite(y[i])) return;  // no usable layout: keep the given order
    x0 = std::min(x0, x[i]);
    x1 = std::max(x1, x[i]);
    y0 = std::min(y0, y[i]);
    y1 = std::max(y1, y[i]);
  }
  const double span = std::max({x1 - x0, y1 - y0, 1e-9});
  std::vector<std::pair<uint64_t, uint32_t>> key(n);
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t qx = (uint32_t)std::min(65535.0, (x[i] - x0) / span * 65535.0);
    const uint32_t qy = (uint32_t)std::min(65535.0, (y[i] - y0) / span * 65535.0);
    key[i] = {hilbert_index(qx, qy), i};
  }
  std::sort(key.begin(), key.end());
  g.identity_order = true;
  for (uint32_t d = 0; d < n; ++d) {
    g.iperm[d] = key[d].second;
    g.perm[key[d].second] = d;
    if (key[d].second != d) g.identity_order = false;
  }
}

// Pinned host arrays of progress counts. A graph takes one on its first call
// and hands it back when destroyed; the arrays are reused, never freed, since
// kernels write them through the device's view of host memory.
std::mutex g_progress_mu;
std::vector<std::pair<uint32_t *, uint32_t *>> g_progress_pool;  // {host, device}

uint32_t *progress_begin(Graph &g, uint64_t settled_on_host) {
  uint32_t *slots = g.progress_slots.load(std::memory_order_acquire);
  if (!slots) {
    std::pair<uint32_t *, uint32_t *> pr{nullptr, nullptr};
    {
      std::lock_guard<std::mutex> lk(g_progress_mu);
      if (!g_progress_pool.empty()) {
        pr = g_progress_pool.back();
        g_progress_pool.pop_back();
      }
    }
    if (!pr.first) {
      void *hp = nullptr, *dp = nullptr;
      hip_check(hipHostMalloc(&hp, 4ull * kProgressSlots, hipHostMallocMapped | hipHostMallocCoherent),
                "hipHostMalloc progress");
      hip_check(hipHostGetDevicePointer(&dp, hp, 0), "hipHostGetDevicePointer");
      pr = {static_cast<uint32_t *>(hp), static_cast<uint32_t *>(dp)};
    }
    g.progress_dev = pr.second;
    slots = pr.first;
  }
  g.progress_host.store(settled_on_host, std::memory_order_relaxed);
  for (uint32_t i = 0; i < kProgressSlots; ++i) __atomic_store_n(&slots[i], 0u, __ATOMIC_RELAXED);
  g.progress_slots.store(slots, std::memory_order_release);
  return g.progress_dev;
}

void progress_finish(Graph &g, uint64_t n) {
  uint64_t counted = 0;
  if (const uint32_t *slots = g.progress_slots.load(std::memory_order_acquire))
    for (uint32_t i = 0; i < kProgressSlots; ++i) counted += __atomic_load_n(&slots[i], __ATOMIC_RELAXED);
  g.progress_host.store(n > counted ? n - counted : 0, std::memory_order_relaxed);
}

DevGraph &device_graph(Graph &g, int device) {
  auto it = g.dev.find(device);
  if (it != g.dev.end()) return *it->second;
  auto dg = std::make_unique<DevGraph>();
  dg->device = device;
  dg->rev_off = dev_upload(g.rev_off);
  dg->fwd_off = dev_upload(g.fwd_off);
  // device node d is original node iperm[d]; its rows stay where the
  // original CSR puts them, with neighbour ids translated (device_edges)
  std::vector<uint4> rows(g.n);
  for (uint32_t d = 0; d < g.n; ++d) {
    const uint32_t v = g.iperm[d];
    rows[d] = make_uint4(g.fwd_off[v], g.fwd_off[v + 1], g.rev_off[v], g.rev_off[v + 1]);
  }
  dg->rows = dev_upload(rows);
  dg->orig = dev_upload(g.iperm);
  dg->perm = dev_upload(g.perm);
  DevGraph &ref = *dg;
  g.dev[device] = std::move(dg);
  return ref;
}

DevGraph *find_device_graph(Graph &g, int device) {
  auto it = g.dev.find(device);
  return it == g.dev.end() ? nullptr : it->second.get();
}

DevEdges *find_device_edges(DevGraph &dg, float speed) {
  for (auto &e : dg.edges)
    if (e->speed == speed) return e.get();
  return nullptr;
}

DevEdges &device_edges(Graph &g, DevGraph &dg, float speed) {
  for (auto &e : dg.edges)
    if (e->speed == speed) return *e;
  auto d = std::make_unique<DevEdges>();
  d->speed = speed;
  const uint32_t m = g.m;
  std::vector<uint2> rev(m), fwd(m);
  std::vector<float2> rtw(m);
  std::vector<float> ftw(m);
  std::vector<uint32_t> reid(m), feid(m);
  float min_es = INFINITY, max_es = 0.0f;
  for (uint32_t v = 0; v < g.n; ++v) {
    for (uint32_t j = g.rev_off[v]; j < g.rev_off[v + 1]; ++j) {
      const uint32_t e = g.rev_eid[j];
      const uint32_t nb = g.esrc[e];
      const float es = edge_seconds(g, nb, v, e, speed, true);
      uint32_t bits;
      std::memcpy(&bits, &es, 4);
      const int32_t tw = g.twin[e];
      // circuit-rank flag of an in-edge: its undirected key is counted here
      // (at the later-settled endpoint) when it or its reverse twin is the
      // key's canonical edge
      const bool count_in = g.canon[e] || (tw >= 0 && g.canon[tw]);
      rev[j] = make_uint2(g.perm[nb] | (count_in ? kCanonBit : 0u), bits);
      rtw[j] = make_float2(tw >= 0 ? g.elen[tw] : NAN, tw >= 0 ? g.eimp[tw] : NAN);
      reid[j] = e;
      if (nb != v) {
        min_es = std::min(min_es, es);
        max_es = std::max(max_es, es);
      }
    }
    for (uint32_t j = g.fwd_off[v]; j < g.fwd_off[v + 1]; ++j) {
      const uint32_t e = g.fwd_eid[j];
      const uint32_t u = g.edst[e];
      const float es = edge_seconds(g, v, u, e, speed, true);
      uint32_t bits;
      std::memcpy(&bits, &es, 4);
      const int32_t tw = g.twin[e];
      // out-edges count only one-way canonical edges (no twin among the in-edges)
      const bool count_out = g.canon[e] && tw < 0;
      fwd[j] = make_uint2(g.perm[u] | (count_out ? kCanonBit : 0u), bits);
      if (count_out) d->has_out_canon = true;
      ftw[j] = tw >= 0 ? g.elen[tw] : NAN;
      feid[j] = e;
    }
  }
  d->min_es = min_es;  // infinite when no edge joins two distinct nodes
  d->max_es = max_es;
  d->rev_adj = dev_upload(rev);
  d->fwd_adj = dev_upload(fwd);
  d->rev_tw = dev_upload(rtw);
  d->fwd_twlen = dev_upload(ftw);
  d->rev_eid = dev_upload(reid);
  d->fwd_eid = dev_upload(feid);
  dg.edges.push_back(std::move(d));
  return *dg.edges.back();
}

// ---------------------------------------------------------------------------
// Threshold conversion (common.rs:98-270), f32 arithmetic, round half away
// from zero like Rust's f32::round.
// ---------------------------------------------------------------------------
namespace {

void betas_from_distances(const uint32_t *d, uint32_t n, float mtw, float *out) {
  if (n == 0) fail(CS_EINVAL, "Input 'distances' cannot be empty.");
  for (uint32_t i = 1; i < n; ++i)
    if (d[i] <= d[i - 1]) fail(CS_EINVAL, "Distances must be unique and sorted in strictly increasing order.");
  for (uint32_t i = 0; i < n; ++i) {
    if (d[i] == 0) fail(CS_EINVAL, "Distances must be positive integers.");
    const float beta = -std::log(mtw) / (float)d[i];
    out[i] = std::round(beta * 100000.0f) / 100000.0f;
  }
}

void distances_from_betas(const float *b, uint32_t n, float mtw, uint32_t *out) {
  if (n == 0) fail(CS_EINVAL, "Input 'betas' cannot be empty.");
  for (uint32_t i = 1; i < n; ++i)
    if (b[i] >= b[i - 1]) fail(CS_EINVAL, "Betas must be unique and sorted in strictly decreasing order.");
  for (uint32_t i = 0; i < n; ++i) {
    if (b[i] <= 0.0f) fail(CS_EINVAL, "Beta values must be greater than zero.");
    const float d = std::round(std::log(mtw) / -b[i]);
    if (d <= 0.0f) fail(CS_EINVAL, "Derived distance must be positive. Check beta values.");
    out[i] = (uint32_t)d;
  }
}

void seconds_from_distances(const uint32_t *d, uint32_t n, float speed, uint32_t *out) {
  if (n == 0) fail(CS_EINVAL, "Input 'distances' cannot be empty.");
  if (speed <= 0.0f) fail(CS_EINVAL, "Speed must be positive.");
  for (uint32_t i = 1; i < n; ++i)
    if (d[i] <= d[i - 1]) fail(CS_EINVAL, "Distances must be unique and sorted in strictly increasing order.");
  for (uint32_t i = 0; i < n; ++i) {
    if (d[i] == 0) fail(CS_EINVAL, "Distances must be positive integers.");
    const float t = std::round((float)d[i] / speed);
    if (t <= 0.0f) fail(CS_EINVAL, "Derived time must be positive. Check distance and speed values.");
    out[i] = (uint32_t)t;
  }
}

void distances_from_seconds(const uint32_t *s, uint32_t n, float speed, uint32_t *out) {
  if (n == 0) fail(CS_EINVAL, "Input 'seconds' cannot be empty.");
  if (speed <= 0.0f) fail(CS_EINVAL, "Speed must be positive.");
  for (uint32_t i = 1; i < n; ++i)
    if (s[i] <= s[i - 1]) fail(CS_EINVAL, "Times must be unique and sorted in strictly increasing order.");
  for (uint32_t i = 0; i < n; ++i) {
    if (s[i] == 0) fail(CS_EINVAL, "Time values must be positive integers.");
    const float d = std::round((float)s[i] * speed);
    if (d <= 0.0f) fail(CS_EINVAL, "Derived distance must be positive. Check time and speed values.");
    out[i] = (uint32_t)d;
  }
}

}  // namespace

int current_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    (void)hipGetLastError();
    fail(CS_EDEVICE, "no HIP device available: the centrality kernels require an AMD GPU (gfx950)");
  }
  int d = 0;
  hip_check(hipGetDevice(&d), "hipGetDevice");
  return d;
}

template <typename F>
int guarded(F &&f) {
  try {
    t_err.clear();
    f();
    return CS_OK;
  } catch (const Error &e) {
    t_err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    t_err = "host allocation failed";
    return CS_ENOMEM;
  } catch (const std::exception &e) {
    t_err = e.what();
    return CS_EINTERNAL;
  }
}

int guarded_call(const std::function<void()> &f) { return guarded(f); }

struct DevBuf {
  void *p = nullptr;
  explicit DevBuf(size_t bytes) { hip_check(hipMalloc(&p, std::max<size_t>(bytes, 4)), "hipMalloc"); }
  ~DevBuf() { (void)hipFree(p); }
  template <typename T>
  T *as() const {
    return static_cast<T *>(p);
  }
};

void check_params(const Graph &g, const cs_params *p) {
  if (!p) fail(CS_EINVAL, "params must not be NULL");
  if (p->n_thresholds == 0) fail(CS_EINVAL, "at least one distance threshold is required");
  if (!p->compute_closeness && !p->compute_betweenness)
    fail(CS_EINVAL,
         "Either or both closeness and betweenness flags is required, but both parameters are False.");
  if (!(p->speed_m_s > 0.0f) || !std::isfinite(p->speed_m_s))
    fail(CS_EINVAL, "speed_m_s must be finite and positive");
  if (!(p->tolerance >= 0.0f && p->tolerance <= 1.0f))
    fail(CS_EINVAL, "Tolerance must be between 0 and 100 (percent), got " + std::to_string(p->tolerance * 100.0f));
  for (uint32_t i = 0; i < p->n_sources; ++i)
    if (p->sources[i] >= g.n)
      fail(CS_EINVAL, "node index " + std::to_string(p->sources[i]) + " does not exist in the graph");
}

CallSpec make_spec(const cs_params *p, uint32_t t0, uint32_t n) {
  if (n == 0) n = p->n_thresholds - t0;
  if (n > kMaxT || t0 + n > p->n_thresholds) fail(CS_EINTERNAL, "threshold chunk out of range");
  CallSpec c;
  c.T = n;
  for (uint32_t i = 0; i < n; ++i) {
    c.dist[i] = (float)p->distances[t0 + i];
    c.beta[i] = p->betas[t0 + i];
    c.seconds[i] = p->seconds[t0 + i];
  }
  // every chunk searches as far as the call's largest threshold, as the
  // reference's single search does (tie bands at a chunk's own cutoff would
  // otherwise differ)
  uint32_t maxs = 0;
  for (uint32_t i = 0; i < p->n_thresholds; ++i) maxs = std::max(maxs, p->seconds[i]);
  c.maxsec = (float)maxs;
  c.speed = p->speed_m_s;
  c.tolerance = p->tolerance;
  c.closeness = p->compute_closeness != 0;
  c.betweenness = p->compute_betweenness != 0;
  c.track_reach = p->track_reach != 0;
  c.ang_unit = p->angular_scaling_unit;
  c.far_offset = p->farness_scaling_offset;
  return c;
}


SourceSet source_set(const Graph &g, const cs_params *p) {
  SourceSet s;
  s.src.reserve(p->n_sources);
  for (uint32_t i = 0; i < p->n_sources; ++i) {
    const uint32_t v = p->sources[i];
    if (!p->eligible[v]) continue;
    const float w = p->source_wt ? p->source_wt[i] : g.weight[v];
    if (std::isnan(w)) continue;
    s.src.push_back(v);
    s.wt.push_back(w);
    s.cyc.push_back(w / g.weight[v]);
  }
  return s;
}

void to_device_sources(const Graph &g, SourceSet &ss) {
  g.order_frozen = true;
  if (g.identity_order) return;
  const size_t n = ss.src.size();
  std::vector<uint32_t> ord(n);
  for (size_t i = 0; i < n; ++i) ord[i] = (uint32_t)i;
  std::stable_sort(ord.begin(), ord.end(),
                   [&](uint32_t a, uint32_t b) { return g.perm[ss.src[a]] < g.perm[ss.src[b]]; });
  SourceSet out;
  out.src.resize(n);
  out.wt.resize(n);
  out.cyc.resize(n);
  for (size_t i = 0; i < n; ++i) {
    out.src[i] = g.perm[ss.src[ord[i]]];
    out.wt[i] = ss.wt[ord[i]];
    out.cyc[i] = ss.cyc[ord[i]];
  }
  ss = std::move(out);
}

void to_device_ids(const Graph &g, std::vector<uint32_t> &ids) {
  g.order_frozen = true;
  if (g.identity_order) return;
  for (auto &v : ids) v = g.perm[v];
}

std::vector<uint8_t> to_device_u8(const Graph &g, const uint8_t *in) {
  std::vector<uint8_t> out(g.n);
  for (uint32_t d = 0; d < g.n; ++d) out[d] = in[g.iperm[d]];
  return out;
}

const uint32_t *result_order(const Graph &g) { return g.identity_order ? nullptr : g.iperm.data(); }

// acc [M][stride][N] (device rows) -> out [M][T][N] (host): drops the padded
// thresholds and puts device node d at original column iperm[d]
void unpack_acc_strided(const double *acc, uint32_t N, uint32_t M, uint32_t T, uint32_t stride, double *out,
                        const uint32_t *iperm, uint32_t out_T, uint32_t t0) {
  if (out_T == 0) out_T = T;
  for (uint32_t m = 0; m < M; ++m)
    for (uint32_t i = 0; i < T; ++i) {
      const double *src = acc + ((size_t)m * stride + i) * N;
      double *dst = out + ((size_t)m * out_T + t0 + i) * N;
      if (!iperm) {
        std::memcpy(dst, src, sizeof(double) * N);
      } else {
        for (uint32_t d = 0; d < N; ++d) dst[iperm[d]] = src[d];
      }
    }
}

void unpack_acc(const std::vector<double> &acc, uint32_t N, uint32_t M, uint32_t T, double *out,
                const uint32_t *iperm) {
  unpack_acc_strided(acc.data(), N, M, T, T, out, iperm);
}

void scale_rows(double *out, uint32_t N, uint32_t T, uint32_t m0, uint32_t m1, double s, uint32_t t0,
                uint32_t n_t) {
  if (s == 1.0) return;
  if (n_t == 0) n_t = T - t0;
  for (uint32_t m = m0; m <= m1; ++m)
    for (uint32_t i = t0; i < t0 + n_t; ++i)
      for (uint32_t v = 0; v < N; ++v) out[((size_t)m * T + i) * N + v] *= s;
}

void check_simplest_graph(const Graph &g) {
  if (!g.is_dual)
    fail(CS_EINVAL,
         "centrality_simplest requires a dual graph for angular analysis. Convert the graph with "
         "cityseer.tools.graphs.nx_to_dual(...) before ingesting it into NetworkStructure.");
  if (!g.slots_ok) fail(CS_EINVAL, g.slots_err);
}

// Predecessor sets are kept over out-row positions (a bitmask up to 64
// out-edges, up to four 16-bit positions beyond); a node with more than four
// tied predecessors at such a wide row is refused by the search itself.
void check_betweenness_graph(const Graph &g) {
  if (g.max_out_deg > 65534)
    fail(CS_EINVAL, "nodes with more than 65,534 out-edges are not supported by the GPU search");
}

void run_shortest_checks(const Graph &g, const cs_params *p) {
  if (p->compute_betweenness) check_betweenness_graph(g);
}


}  // namespace cs

using namespace cs;

// Host side of a batch: each job's sources prepared as cs_centrality_shortest
// prepares them, device buffers built on the first run.
struct cs_batch {
  struct JobHost {
    cs_graph *h = nullptr;
    CallSpec spec;
    SourceSet ss;
    std::vector<uint8_t> elig;
    double *out = nullptr;
    uint32_t *reach = nullptr;
    bool betw = false;
    double betw_scale = 1.0;
    size_t acc_off = 0;
    uint32_t out_T = 0, t0 = 0;  // the request's thresholds; this job's first (a chunk of at most kMaxT)
  };
  int device = -1;
  std::vector<JobHost> jobs;
  int n_requests = 0;  // cs_batch_add_shortest calls; a request runs as one job per threshold chunk
  uint32_t NT = 1;
  Tuning tune;
  hipStream_t stream = nullptr;
  std::vector<void *> bufs;
  double *d_acc = nullptr;
  size_t acc_words = 0;
  uint32_t *d_reach = nullptr;
  ShortestBatch *sb = nullptr;
  bool have_results = false;

  void drop_device() {
    if (sb) batch_delete(sb);
    sb = nullptr;
    for (void *q : bufs) (void)hipFree(q);
    bufs.clear();
    if (d_acc) (void)hipFree(d_acc);
    if (d_reach) (void)hipFree(d_reach);
    d_acc = nullptr;
    d_reach = nullptr;
    have_results = false;
  }
  ~cs_batch() {
    drop_device();
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename T>
  T *upload(const T *src, size_t n) {
    void *q = nullptr;
    hip_check(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc batch");
    bufs.push_back(q);
    if (n) hip_check(hipMemcpy(q, src, sizeof(T) * n, hipMemcpyHostToDevice), "upload batch");
    return static_cast<T *>(q);
  }
  void build();
};


extern "C" {

const char *cs_last_error(void) { return t_err.c_str(); }

const char *cs_version(void) { return "cityseer-amd 0.1.0 (gfx950)"; }

int cs_pair_distances_betas_time(float speed, int mode, const uint32_t *distances, const float *betas,
                                 const float *minutes, uint32_t n, float mtw, uint32_t *od, float *ob,
                                 uint32_t *os) {
  return guarded([&] {
    if (mode == 0) {
      for (uint32_t i = 0; i < n; ++i) od[i] = distances[i];
      betas_from_distances(od, n, mtw, ob);
      seconds_from_distances(od, n, speed, os);
    } else if (mode == 1) {
      for (uint32_t i = 0; i < n; ++i) ob[i] = betas[i];
      distances_from_betas(ob, n, mtw, od);
      seconds_from_distances(od, n, speed, os);
    } else if (mode == 2) {
      for (uint32_t i = 0; i < n; ++i) os[i] = (uint32_t)std::round(minutes[i] * 60.0f);
      distances_from_seconds(os, n, speed, od);
      betas_from_distances(od, n, mtw, ob);
    } else {
      fail(CS_EINVAL,
           "Please provide exactly one of the following arguments: 'distances', 'betas', or 'minutes'.");
    }
  });
}

int cs_distances_from_betas(const float *betas, uint32_t n, float mtw, uint32_t *out) {
  return guarded([&] { distances_from_betas(betas, n, mtw, out); });
}
int cs_betas_from_distances(const uint32_t *d, uint32_t n, float mtw, float *out) {
  return guarded([&] { betas_from_distances(d, n, mtw, out); });
}
int cs_distances_from_seconds(const uint32_t *s, uint32_t n, float speed, uint32_t *out) {
  return guarded([&] { distances_from_seconds(s, n, speed, out); });
}
int cs_seconds_from_distances(const uint32_t *d, uint32_t n, float speed, uint32_t *out) {
  return guarded([&] { seconds_from_distances(d, n, speed, out); });
}

int cs_graph_create(uint32_t n_nodes, const double *z, const uint8_t *live, const float *weight,
                    uint32_t n_edges, const uint32_t *src, const uint32_t *dst, const uint32_t *edge_key,
                    const float *length, const float *imp_factor, const float *angle_sum,
                    const float *seconds, const int32_t *primal_key, int is_dual, cs_graph **out) {
  return guarded([&] {
    if (!out) fail(CS_EINVAL, "out must not be NULL");
    if (n_nodes >= kNodeMask) fail(CS_EINVAL, "too many nodes");
    if (n_edges && (!src || !dst || !length)) fail(CS_EINVAL, "edge arrays must not be NULL");
    auto h = std::make_unique<cs_graph>();
    Graph &g = h->g;
    g.n = n_nodes;
    g.m = n_edges;
    g.is_dual = is_dual != 0;
    g.z.resize(n_nodes);
    g.live.resize(n_nodes);
    g.weight.resize(n_nodes);
    for (uint32_t i = 0; i < n_nodes; ++i) {
      g.z[i] = z ? z[i] : NAN;
      g.live[i] = live ? live[i] : 1;
      g.weight[i] = weight ? weight[i] : 1.0f;
    }
    g.esrc.assign(src, src + n_edges);
    g.edst.assign(dst, dst + n_edges);
    g.ekey.resize(n_edges);
    g.elen.assign(length, length + n_edges);
    g.eimp.resize(n_edges);
    g.eang.resize(n_edges);
    g.esec.resize(n_edges);
    g.epk.resize(n_edges);
    for (uint32_t e = 0; e < n_edges; ++e) {
      if (src[e] >= n_nodes || dst[e] >= n_nodes)
        fail(CS_EINVAL, "edge " + std::to_string(e) + " references a node that does not exist");
      g.ekey[e] = edge_key ? edge_key[e] : 0;
      g.eimp[e] = imp_factor ? imp_factor[e] : 1.0f;
      g.eang[e] = angle_sum ? angle_sum[e] : 0.0f;
      g.esec[e] = seconds ? seconds[e] : NAN;
      g.epk[e] = primal_key ? primal_key[e] : -1;
      if (!(g.eimp[e] > 0.0f) || !std::isfinite(g.eimp[e]))
        fail(CS_EINVAL, "Invalid impedance factor for edge " + std::to_string(e) +
                            ": impedance must be finite and positive (> 0.0).");
    }
    build_topology(g);
    g.perm.resize(n_nodes);
    g.iperm.resize(n_nodes);
    for (uint32_t i = 0; i < n_nodes; ++i) g.perm[i] = g.iperm[i] = i;
    *out = h.release();
  });
}

void cs_graph_destroy(cs_graph *h) {
  if (!h) return;
  if (uint32_t *slots = h->g.progress_slots.load()) {
    std::lock_guard<std::mutex> lk(g_progress_mu);
    g_progress_pool.emplace_back(slots, h->g.progress_dev);
  }
  for (auto &kv : h->g.dev) {
    int prev = 0;
    if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(kv.first) == hipSuccess) {
      (void)hipFree(kv.second->rev_off);
      (void)hipFree(kv.second->fwd_off);
      (void)hipFree(kv.second->rows);
      (void)hipFree(kv.second->orig);
      (void)hipFree(kv.second->perm);
      (void)hipFree(kv.second->xbuf);
      for (auto &e : kv.second->edges) free_edges(*e);
      (void)hipSetDevice(prev);
    }
  }
  (void)hipGetLastError();
  delete h;
}

int cs_graph_set_xy(cs_graph *h, const double *x, const double *y) {
  return guarded([&] {
    if (!h || !x || !y) fail(CS_EINVAL, "graph, x and y must not be NULL");
    std::lock_guard<std::mutex> lk(h->g.mu);
    set_node_order_from_xy(h->g, x, y);
  });
}

int cs_graph_upload(cs_graph *h, int device, float speed_m_s) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    (void)current_device();  // raises without a GPU
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
      fail(CS_EINVAL, "device " + std::to_string(device) + " does not exist (" + std::to_string(count) + " visible)");
    if (speed_m_s != 0.0f && !(std::isfinite(speed_m_s) && speed_m_s > 0.0f))
      fail(CS_EINVAL, "speed_m_s must be finite and positive (or 0: adjacency only)");
    int prev = 0;
    hip_check(hipGetDevice(&prev), "hipGetDevice");
    hip_check(hipSetDevice(device), "hipSetDevice");
    struct Restore {
      int d;
      ~Restore() { (void)hipSetDevice(d); }
    } restore{prev};
    Graph &g = h->g;
    std::lock_guard<std::mutex> lk(g.mu);
    g.order_frozen = true;
    DevGraph &dg = device_graph(g, device);
    if (speed_m_s > 0.0f) (void)device_edges(g, dg, speed_m_s);
    hip_check(hipDeviceSynchronize(), "upload");
  });
}

int cs_graph_node_order(const cs_graph *h, uint32_t *order) {
  return guarded([&] {
    if (!h || !order) fail(CS_EINVAL, "graph and order must not be NULL");
    std::copy(h->g.iperm.begin(), h->g.iperm.end(), order);
  });
}

int cs_graph_info(const cs_graph *h, uint32_t *n, uint32_t *m, uint32_t *mo, uint32_t *mi) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    if (n) *n = h->g.n;
    if (m) *m = h->g.m;
    if (mo) *mo = h->g.max_out_deg;
    if (mi) *mi = h->g.max_in_deg;
  });
}

uint64_t cs_graph_progress(const cs_graph *h) {
  if (!h) return 0;
  const Graph &g = h->g;
  uint64_t n = g.progress_host.load(std::memory_order_relaxed);
  if (const uint32_t *slots = g.progress_slots.load(std::memory_order_acquire))
    for (uint32_t i = 0; i < kProgressSlots; ++i) n += __atomic_load_n(&slots[i], __ATOMIC_RELAXED);
  return n;
}

int cs_last_stats(cs_stats *out) {
  if (!out) return CS_EINVAL;
  *out = t_stats;
  return CS_OK;
}

int cs_centrality_shortest(cs_graph *h, const cs_params *p, double *out, uint32_t *reach_totals) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h || !out) fail(CS_EINVAL, "graph and out must not be NULL");
    Graph &g = h->g;
    check_params(g, p);
    const int dev = current_device();
    const CallSpec spec = make_spec(p, 0, chunk_len(p->n_thresholds, 0));
    SourceSet ss = source_set(g, p);
    to_device_sources(g, ss);
    const std::vector<uint8_t> elig = to_device_u8(g, p->eligible);
    const uint32_t N = g.n, T = p->n_thresholds, M = 7;
    OwnedStream stream;
    DevBuf acc(sizeof(double) * (size_t)N * M * T), reach(4 * T), d_src(4 * ss.src.size()),
        d_wt(4 * ss.src.size()), d_cyc(4 * ss.src.size()), d_elig(N);
    hip_check(hipMemsetAsync(acc.p, 0, sizeof(double) * (size_t)N * M * T, stream), "memset");
    hip_check(hipMemsetAsync(reach.p, 0, 4 * T, stream), "memset");
    if (!ss.src.empty()) {
      hip_check(hipMemcpyAsync(d_src.p, ss.src.data(), 4 * ss.src.size(), hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_wt.p, ss.wt.data(), 4 * ss.wt.size(), hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_cyc.p, ss.cyc.data(), 4 * ss.cyc.size(), hipMemcpyHostToDevice, stream), "copy");
    }
    hip_check(hipMemcpyAsync(d_elig.p, elig.data(), N, hipMemcpyHostToDevice, stream), "copy");
    {
      std::lock_guard<std::mutex> lk(g.mu);
      DevGraph &dg = device_graph(g, dev);
      DevEdges &de = device_edges(g, dg, spec.speed);
      run_shortest_checks(g, p);
      uint32_t *prog = progress_begin(g, p->n_sources - ss.src.size());
      // a call of several threshold chunks reports progress in its last one
      for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
        launch_shortest(g, dg, de, make_spec(p, t0, chunk_len(T, t0)), d_src.as<uint32_t>(), d_wt.as<float>(),
                        d_cyc.as<float>(), (uint32_t)ss.src.size(), d_elig.as<uint8_t>(),
                        acc.as<double>() + (size_t)t0 * N, reach.as<uint32_t>() + t0, stream, t_stats, T,
                        t0 + kMaxT >= T ? prog : nullptr);
      progress_finish(g, p->n_sources);
    }
    std::vector<double> hacc((size_t)N * M * T);
    std::vector<uint32_t> hreach(T);
    hip_check(hipMemcpyAsync(hacc.data(), acc.p, sizeof(double) * hacc.size(), hipMemcpyDeviceToHost, stream),
              "copy out");
    hip_check(hipMemcpyAsync(hreach.data(), reach.p, 4 * T, hipMemcpyDeviceToHost, stream), "copy out");
    hip_check(hipStreamSynchronize(stream), "sync");
    unpack_acc(hacc, N, M, T, out, result_order(g));
    if (p->compute_betweenness) scale_rows(out, N, T, 5, 6, p->betw_scale);
    if (reach_totals)
      for (uint32_t i = 0; i < T; ++i) reach_totals[i] = hreach[i];
    t_stats.sources = ss.src.size();
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_centrality_shortest_device(cs_graph *h, const cs_params *p, const float *d_cycles_wt, double *d_acc,
                                  uint32_t *d_reach, void *stream_ptr) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h || !p || !d_acc) fail(CS_EINVAL, "graph, params and d_acc must not be NULL");
    Graph &g = h->g;
    if (p->n_thresholds == 0) fail(CS_EINVAL, "at least one distance threshold is required");
    if (!p->compute_closeness && !p->compute_betweenness)
      fail(CS_EINVAL, "Either or both closeness and betweenness flags is required, but both parameters are False.");
    if (!p->sources || !p->source_wt || !d_cycles_wt || !p->eligible)
      fail(CS_EINVAL, "device sources, weights, cycle weights and eligibility must not be NULL");
    const int dev = current_device();
    const uint32_t T = p->n_thresholds, N = g.n;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, dev);
    DevEdges &de = device_edges(g, dg, p->speed_m_s);
    run_shortest_checks(g, p);
    uint32_t *prog = progress_begin(g, 0);
    for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
      launch_shortest_original(g, dg, de, make_spec(p, t0, chunk_len(T, t0)), p->sources, p->source_wt,
                               d_cycles_wt, p->n_sources, p->eligible, d_acc + (size_t)t0 * N,
                               d_reach ? d_reach + t0 : nullptr, stream, t_stats, T,
                               t0 + kMaxT >= T ? prog : nullptr);
    progress_finish(g, p->n_sources);
    t_stats.sources = p->n_sources;
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_betweenness_od_shortest(cs_graph *h, const cs_params *p, uint32_t n_pairs, const uint32_t *origins,
                               const uint32_t *destinations, const float *weights, double *out) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h || !p || !out) fail(CS_EINVAL, "graph, params and out must not be NULL");
    if (n_pairs && (!origins || !destinations || !weights))
      fail(CS_EINVAL, "origins, destinations and weights must not be NULL");
    Graph &g = h->g;
    if (p->n_thresholds == 0) fail(CS_EINVAL, "at least one distance threshold is required");
    if (!(p->speed_m_s > 0.0f) || !std::isfinite(p->speed_m_s))
      fail(CS_EINVAL, "speed_m_s must be finite and positive");
    if (!(p->tolerance >= 0.0f && p->tolerance <= 1.0f))
      fail(CS_EINVAL, "Tolerance must be between 0 and 100 (percent), got " + std::to_string(p->tolerance * 100.0f));
    const uint32_t N = g.n;
    for (uint32_t i = 0; i < n_pairs; ++i)
      if (origins[i] >= N || destinations[i] >= N)
        fail(CS_EINVAL, "node index " + std::to_string(std::max(origins[i], destinations[i])) +
                            " does not exist in the graph");
    // OdMatrix::new (centrality.rs:61-82): one row per origin, a later pair
    // overwriting an earlier one with the same destination
    std::vector<uint32_t> order(n_pairs);
    for (uint32_t i = 0; i < n_pairs; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return origins[a] != origins[b] ? origins[a] < origins[b] : destinations[a] < destinations[b];
    });
    std::vector<uint32_t> src, off{0}, dest;
    std::vector<float> w;
    for (uint32_t k = 0; k < n_pairs;) {
      const uint32_t o = origins[order[k]];
      const bool live = g.live[o] != 0;  // betweenness_od_shortest skips non-live origins (:2484-2486)
      while (k < n_pairs && origins[order[k]] == o) {
        uint32_t last = k;
        while (last + 1 < n_pairs && origins[order[last + 1]] == o &&
               destinations[order[last + 1]] == destinations[order[k]])
          ++last;
        if (live) {
          dest.push_back(destinations[order[last]]);
          w.push_back(weights[order[last]]);
        }
        k = last + 1;
      }
      if (live) {
        src.push_back(o);
        off.push_back((uint32_t)dest.size());
      }
    }
    if (!g.identity_order) {
      // device node ids, origins ascending in device order
      std::vector<uint32_t> ord(src.size());
      for (uint32_t i = 0; i < ord.size(); ++i) ord[i] = i;
      std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return g.perm[src[a]] < g.perm[src[b]]; });
      std::vector<uint32_t> src2, off2{0}, dest2;
      std::vector<float> w2;
      for (const uint32_t o : ord) {
        src2.push_back(g.perm[src[o]]);
        for (uint32_t k = off[o]; k < off[o + 1]; ++k) {
          dest2.push_back(g.perm[dest[k]]);
          w2.push_back(w[k]);
        }
        off2.push_back((uint32_t)dest2.size());
      }
      src.swap(src2);
      off.swap(off2);
      dest.swap(dest2);
      w.swap(w2);
    }
    const uint32_t T = p->n_thresholds, M = 7;
    const uint32_t ns = (uint32_t)src.size();
    const int dev = current_device();
    OwnedStream stream;
    DevBuf acc(sizeof(double) * (size_t)N * M * T), d_src(4 * std::max<size_t>(ns, 1)),
        d_wt(4 * std::max<size_t>(ns, 1)), d_elig(N), d_off(4 * off.size()), d_dest(4 * std::max<size_t>(dest.size(), 1)),
        d_w(4 * std::max<size_t>(w.size(), 1));
    const std::vector<float> ones(std::max<size_t>(ns, 1), 1.0f);
    hip_check(hipMemsetAsync(acc.p, 0, sizeof(double) * (size_t)N * M * T, stream), "memset");
    hip_check(hipMemsetAsync(d_elig.p, 0, N, stream), "memset");
    if (ns) {
      hip_check(hipMemcpyAsync(d_src.p, src.data(), 4 * ns, hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_wt.p, ones.data(), 4 * ns, hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_off.p, off.data(), 4 * off.size(), hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_dest.p, dest.data(), 4 * dest.size(), hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_w.p, w.data(), 4 * w.size(), hipMemcpyHostToDevice, stream), "copy");
      std::lock_guard<std::mutex> lk(g.mu);
      DevGraph &dg = device_graph(g, dev);
      DevEdges &de = device_edges(g, dg, p->speed_m_s);
      check_betweenness_graph(g);
      uint32_t *prog = progress_begin(g, N - ns);
      for (uint32_t t0 = 0; t0 < T; t0 += kMaxT) {
        CallSpec spec = make_spec(p, t0, chunk_len(T, t0));
        spec.closeness = false;
        spec.betweenness = true;
        spec.track_reach = false;
        JobSpec js{&g, &dg, &de, spec, d_src.as<uint32_t>(), d_wt.as<float>(), d_wt.as<float>(), ns,
                   d_elig.as<uint8_t>(), acc.as<double>() + (size_t)t0 * N, nullptr, default_tuning().delta_scale};
        js.tstride = T;
        js.d_od_off = d_off.as<uint32_t>();
        js.d_od_dest = d_dest.as<uint32_t>();
        js.d_od_w = d_w.as<float>();
        launch_shortest_batch({js}, spec.T, default_tuning(), stream, t_stats, nullptr,
                              t0 + kMaxT >= T ? prog : nullptr);
      }
    }
    std::vector<double> hacc((size_t)N * M * T);
    hip_check(hipMemcpyAsync(hacc.data(), acc.p, sizeof(double) * hacc.size(), hipMemcpyDeviceToHost, stream),
              "copy out");
    hip_check(hipStreamSynchronize(stream), "sync");
    // rows 5 and 6 (betweenness, betweenness_beta) of the [M][T][N] accumulators
    unpack_acc_strided(hacc.data() + (size_t)5 * T * N, N, 2, T, T, out, result_order(g));
    t_stats.sources = ns;
    progress_finish(g, N);
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_segment_centrality(cs_graph *h, const cs_params *p, double *out) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h || !out) fail(CS_EINVAL, "graph and out must not be NULL");
    Graph &g = h->g;
    check_params(g, p);
    const int dev = current_device();
    const uint32_t N = g.n, T = p->n_thresholds, M = 4;
    // the reference looks up twin edges (graph.rs:1278-1292) and panics if one is missing
    for (uint32_t e = 0; e < g.m; ++e)
      if (g.twin[e] < 0)
        fail(CS_EINVAL, "segment centrality needs every edge to have a reverse edge with the same key (edge " +
                            std::to_string(e) + ")");
    std::vector<uint32_t> srcs;
    for (uint32_t i = 0; i < p->n_sources; ++i)
      if (g.live[p->sources[i]]) srcs.push_back(p->sources[i]);
    to_device_ids(g, srcs);
    std::sort(srcs.begin(), srcs.end());
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, dev);
    DevEdges &de = device_edges(g, dg, p->speed_m_s);
    OwnedStream stream;
    DevBuf acc(sizeof(double) * (size_t)N * M * T), d_src(4 * srcs.size());
    hip_check(hipMemsetAsync(acc.p, 0, sizeof(double) * (size_t)N * M * T, stream), "memset");
    if (!srcs.empty())
      hip_check(hipMemcpyAsync(d_src.p, srcs.data(), 4 * srcs.size(), hipMemcpyHostToDevice, stream), "copy");
    uint32_t *prog = progress_begin(g, p->n_sources - srcs.size());
    for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
      launch_segment(g, dg, de, make_spec(p, t0, chunk_len(T, t0)), d_src.as<uint32_t>(), (uint32_t)srcs.size(),
                     acc.as<double>() + (size_t)t0 * N, stream, t_stats, T, t0 + kMaxT >= T ? prog : nullptr);
    std::vector<double> hacc((size_t)N * M * T);
    hip_check(hipMemcpyAsync(hacc.data(), acc.p, sizeof(double) * hacc.size(), hipMemcpyDeviceToHost, stream),
              "copy out");
    hip_check(hipStreamSynchronize(stream), "sync");
    unpack_acc(hacc, N, M, T, out, result_order(g));
    t_stats.sources = srcs.size();
    progress_finish(g, p->n_sources);
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_centrality_simplest(cs_graph *h, const cs_params *p, double *out, uint32_t *reach_totals) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h || !out) fail(CS_EINVAL, "graph and out must not be NULL");
    Graph &g = h->g;
    check_simplest_graph(g);
    check_params(g, p);
    const int dev = current_device();
    const SourceSet ss = source_set(g, p);
    const uint32_t N = g.n, T = p->n_thresholds, M = 4;
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, dev);
    DevEdges &de = device_edges(g, dg, p->speed_m_s);
    OwnedStream stream;
    DevBuf acc(sizeof(double) * (size_t)N * M * T), reach(4 * T), d_src(4 * ss.src.size()),
        d_wt(4 * ss.src.size()), d_elig(N);
    hip_check(hipMemsetAsync(acc.p, 0, sizeof(double) * (size_t)N * M * T, stream), "memset");
    hip_check(hipMemsetAsync(reach.p, 0, 4 * T, stream), "memset");
    if (!ss.src.empty()) {
      hip_check(hipMemcpyAsync(d_src.p, ss.src.data(), 4 * ss.src.size(), hipMemcpyHostToDevice, stream), "copy");
      hip_check(hipMemcpyAsync(d_wt.p, ss.wt.data(), 4 * ss.wt.size(), hipMemcpyHostToDevice, stream), "copy");
    }
    hip_check(hipMemcpyAsync(d_elig.p, p->eligible, N, hipMemcpyHostToDevice, stream), "copy");
    uint32_t *prog = progress_begin(g, p->n_sources - ss.src.size());
    for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
      launch_simplest(g, dg, de, make_spec(p, t0, chunk_len(T, t0)), d_src.as<uint32_t>(), d_wt.as<float>(),
                      (uint32_t)ss.src.size(), d_elig.as<uint8_t>(), acc.as<double>() + (size_t)t0 * N,
                      reach.as<uint32_t>() + t0, stream, t_stats, T, t0 + kMaxT >= T ? prog : nullptr);
    std::vector<double> hacc((size_t)N * M * T);
    std::vector<uint32_t> hreach(T);
    hip_check(hipMemcpyAsync(hacc.data(), acc.p, sizeof(double) * hacc.size(), hipMemcpyDeviceToHost, stream),
              "copy out");
    hip_check(hipMemcpyAsync(hreach.data(), reach.p, 4 * T, hipMemcpyDeviceToHost, stream), "copy out");
    hip_check(hipStreamSynchronize(stream), "sync");
    unpack_acc(hacc, N, M, T, out, nullptr);
    if (p->compute_betweenness) scale_rows(out, N, T, 3, 3, p->betw_scale);
    if (reach_totals)
      for (uint32_t i = 0; i < T; ++i) reach_totals[i] = hreach[i];
    t_stats.sources = ss.src.size();
    progress_finish(g, p->n_sources);
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

}  // extern "C"

namespace {

// Single-source search with device outputs copied back.
struct Single {
  std::vector<float> secs;
  std::vector<int32_t> pred, last;
  std::vector<double> sigma;
  std::vector<uint32_t> preds, npreds;
};

Single run_single(cs_graph *h, uint32_t src, uint32_t max_seconds, float speed, float tolerance, int mode,
                  uint32_t max_preds) {
  Graph &g = h->g;
  if (src >= g.n)
    fail(CS_EINVAL, "src_idx " + std::to_string(src) + " out of range for network with node_bound " +
                        std::to_string(g.n));
  if (!std::isfinite(speed) || !(speed > 0.0f)) fail(CS_EINVAL, "speed_m_s must be finite and positive");
  if (mode == 1) check_betweenness_graph(g);
  const int dev = current_device();
  const uint32_t N = g.n;
  std::lock_guard<std::mutex> lk(g.mu);
  DevGraph &dg = device_graph(g, dev);
  DevEdges &de = device_edges(g, dg, speed);
  OwnedStream stream;
  Single s;
  s.secs.assign(N, INFINITY);
  s.pred.assign(N, -1);
  s.last.assign(N, -1);
  DevBuf d_secs(4ull * N), d_pred(4ull * N), d_last(4ull * N), d_sigma(8ull * N), d_np(4ull * N),
      d_pl(4ull * N * std::max<uint32_t>(max_preds, 1));
  hip_check(hipMemcpyAsync(d_secs.p, s.secs.data(), 4ull * N, hipMemcpyHostToDevice, stream), "copy");
  hip_check(hipMemsetAsync(d_pred.p, 0xff, 4ull * N, stream), "memset");
  hip_check(hipMemsetAsync(d_last.p, 0xff, 4ull * N, stream), "memset");
  hip_check(hipMemsetAsync(d_sigma.p, 0, 8ull * N, stream), "memset");
  hip_check(hipMemsetAsync(d_np.p, 0, 4ull * N, stream), "memset");
  TreeOut to;
  to.secs = d_secs.as<float>();
  to.pred = d_pred.as<int32_t>();
  to.last_eid = d_last.as<int32_t>();
  to.sigma = d_sigma.as<double>();
  to.pred_list = d_pl.as<uint32_t>();
  to.n_preds = d_np.as<uint32_t>();
  to.max_preds = max_preds;
  launch_single(g, dg, de, g.perm[src], (float)max_seconds, tolerance, mode, to, stream);
  hip_check(hipMemcpyAsync(s.secs.data(), d_secs.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
  hip_check(hipMemcpyAsync(s.pred.data(), d_pred.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
  hip_check(hipMemcpyAsync(s.last.data(), d_last.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
  if (mode == 1) {
    s.sigma.resize(N);
    s.npreds.resize(N);
    s.preds.resize((size_t)N * max_preds);
    hip_check(hipMemcpyAsync(s.sigma.data(), d_sigma.p, 8ull * N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(s.npreds.data(), d_np.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
    if (max_preds)
      hip_check(hipMemcpyAsync(s.preds.data(), d_pl.p, 4ull * N * max_preds, hipMemcpyDeviceToHost, stream),
                "copy");
  }
  hip_check(hipStreamSynchronize(stream), "sync");
  if (!g.identity_order) {
    // device node order -> original order (edge ids are original already)
    const std::vector<float> secs = s.secs;
    const std::vector<int32_t> pred = s.pred, last = s.last;
    for (uint32_t v = 0; v < N; ++v) {
      const uint32_t d = g.perm[v];
      s.secs[v] = secs[d];
      s.pred[v] = pred[d] >= 0 ? (int32_t)g.iperm[pred[d]] : -1;
      s.last[v] = last[d];
    }
    if (mode == 1) {
      const std::vector<double> sigma = s.sigma;
      const std::vector<uint32_t> np = s.npreds, pl = s.preds;
      for (uint32_t v = 0; v < N; ++v) {
        const uint32_t d = g.perm[v];
        s.sigma[v] = sigma[d];
        s.npreds[v] = np[d];
        const uint32_t k = std::min(np[d], max_preds);
        for (uint32_t i = 0; i < k; ++i) s.preds[(size_t)v * max_preds + i] = g.iperm[pl[(size_t)d * max_preds + i]];
      }
    }
  }
  return s;
}

// visited order = reference settle order: ascending (seconds, node)
uint32_t settle_order(const std::vector<float> &secs, uint32_t *order) {
  std::vector<uint32_t> idx;
  for (uint32_t i = 0; i < secs.size(); ++i)
    if (std::isfinite(secs[i])) idx.push_back(i);
  std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
    return secs[a] < secs[b] || (secs[a] == secs[b] && a < b);
  });
  if (order) std::copy(idx.begin(), idx.end(), order);
  return (uint32_t)idx.size();
}

}  // namespace

extern "C" {

// Accessibility rows (out_count / out_weighted / out_nearest) or, with out_div,
// the diversity rows of the same per-source class counts instead.
// Numeric maps of a statistics call: values [n_maps][n_entries] (NaN: none),
// out [n_maps][kStatMetrics][T][N].
struct StatsIn {
  uint32_t n_maps = 0;
  const float *values = nullptr;
  double *out = nullptr;
};

static int access_impl(cs_graph *h, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                       uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node,
                       const uint32_t *assign_entry, const double *assign_dist, uint32_t n_thresholds,
                       const uint32_t *distances, const float *betas, const float *max_curve_wts,
                       uint32_t max_seconds, float speed_m_s, int angular, double *out_count,
                       double *out_weighted, double *out_nearest, double *out_div,
                       const StatsIn *stats_in = nullptr) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    Graph &g = h->g;
    const uint32_t N = g.n, T = n_thresholds, C = n_classes;
    if (T == 0) fail(CS_EINVAL, "at least one distance is required");
    if (!distances || !betas || !max_curve_wts) fail(CS_EINVAL, "distances, betas and max_curve_wts must not be NULL");
    if (n_entries && (!entry_group || (!entry_class && !stats_in)))
      fail(CS_EINVAL, "entry_group and entry_class must not be NULL");
    if (stats_in && stats_in->n_maps && n_entries && !stats_in->values) fail(CS_EINVAL, "values must not be NULL");
    if (n_assign && (!assign_node || !assign_entry || !assign_dist))
      fail(CS_EINVAL, "assign_node, assign_entry and assign_dist must not be NULL");
    if (!out_div && !stats_in && (size_t)C * N && (!out_count || !out_weighted || !out_nearest))
      fail(CS_EINVAL, "outputs must not be NULL");
    if (!std::isfinite(speed_m_s) || !(speed_m_s > 0.0f)) fail(CS_EINVAL, "speed_m_s must be finite and positive");
    if (angular) {
      // validate_dual_for_angular (graph.rs:1333-1341), then the endpoint slots the search needs
      if (!g.is_dual)
        fail(CS_EINVAL, std::string(out_div ? "mixed_uses" : stats_in ? "stats" : "accessibility") +
                            " requires a dual graph for angular analysis. Convert the graph with "
                            "cityseer.tools.graphs.nx_to_dual(...) before ingesting it into NetworkStructure.");
      if (!g.slots_ok) fail(CS_EINVAL, g.slots_err);
    }
    for (uint32_t t = 0; t < T; ++t)
      if (!(max_curve_wts[t] >= 0.0f && max_curve_wts[t] <= 1.0f) || max_curve_wts[t] == 0.0f)
        fail(CS_EINVAL, "Max curve weight must be in the range [0, 1].");
    for (uint32_t e = 0; e < n_entries; ++e) {
      if (entry_group[e] >= n_entries) fail(CS_EINVAL, "entry_group values must be below n_entries");
      if (!stats_in && (entry_class[e] < -1 || entry_class[e] >= (int32_t)C))
        fail(CS_EINVAL, "entry_class out of range");
    }
    for (uint32_t i = 0; i < n_assign; ++i) {
      if (assign_node[i] >= N) fail(CS_EINVAL, "assignment node out of range");
      if (assign_entry[i] >= n_entries) fail(CS_EINVAL, "assignment entry out of range");
      if (!(assign_dist[i] >= 0.0) || !std::isfinite(assign_dist[i]))
        fail(CS_EINVAL, "assignment distances must be finite and non-negative");
    }
    const size_t CTN = (size_t)C * T * N;
    const uint32_t SM = stats_in ? stats_in->n_maps * kStatMetrics : 0;  // statistics rows per threshold
    std::vector<double> stat_init;
    if (stats_in) {
      // StatsResult::new (data.rs:262-295): sums and counts start at 0, the rest at NaN
      stat_init.assign((size_t)SM * T * N, std::nan(""));
      for (uint32_t mp = 0; mp < stats_in->n_maps; ++mp)
        for (uint32_t m : {0u, 1u, 6u, 7u})
          std::fill_n(stat_init.begin() + ((size_t)mp * kStatMetrics + m) * T * N, (size_t)T * N, 0.0);
      std::copy(stat_init.begin(), stat_init.end(), stats_in->out);
    } else if (out_div) {
      std::fill(out_div, out_div + (size_t)kDivMetrics * T * N, 0.0);
    } else {
      std::fill(out_count, out_count + CTN, 0.0);
      std::fill(out_weighted, out_weighted + CTN, 0.0);
      std::fill(out_nearest, out_nearest + (size_t)C * N, std::nan(""));
    }
    if (N == 0 || n_assign == 0 || (stats_in ? stats_in->n_maps == 0 : C == 0)) return;
    // members: assignments grouped by dedupe group, assignment order within a
    // group; groups of two or more members get a slot of the per-wave tables
    std::vector<uint32_t> gcount(n_entries + 1, 0);
    for (uint32_t i = 0; i < n_assign; ++i) ++gcount[entry_group[assign_entry[i]] + 1];
    std::vector<uint32_t> gslot(n_entries, kSoloGroup);
    uint32_t n_shared = 0, n_groups = 0;
    for (uint32_t x = 0; x < n_entries; ++x) {
      if (gcount[x + 1] > 1) gslot[x] = n_shared++;
      if (gcount[x + 1] > 0) ++n_groups;
    }
    for (uint32_t x = 0; x < n_entries; ++x) gcount[x + 1] += gcount[x];
    std::vector<uint4> mem(n_assign);
    std::vector<uint32_t> mem_node(n_assign);
    for (uint32_t i = 0; i < n_assign; ++i) {
      const uint32_t e = assign_entry[i], grp = entry_group[e];
      const uint32_t m = gcount[grp]++;
      const uint32_t node = g.identity_order ? assign_node[i] : g.perm[assign_node[i]];
      const float dd = (float)assign_dist[i];  // data.rs:548 casts the f64 distance to f32
      uint32_t bits;
      std::memcpy(&bits, &dd, 4);
      mem[m] = make_uint4(node, e, bits, gslot[grp]);
      mem_node[m] = node;
    }
    // members of each device node
    std::vector<uint32_t> node_off(N + 1, 0), node_mem(n_assign);
    for (uint32_t m = 0; m < n_assign; ++m) ++node_off[mem_node[m] + 1];
    for (uint32_t v = 0; v < N; ++v) node_off[v + 1] += node_off[v];
    {
      std::vector<uint32_t> at(node_off.begin(), node_off.end() - 1);
      for (uint32_t m = 0; m < n_assign; ++m) node_mem[at[mem_node[m]]++] = m;
    }
    // sources: every live node (data.rs:681-686)
    std::vector<uint32_t> srcs;
    for (uint32_t v = 0; v < N; ++v)
      if (g.live[v]) srcs.push_back(v);
    to_device_ids(g, srcs);
    std::sort(srcs.begin(), srcs.end());
    AccessOut ao;
    ao.N = N;
    ao.tstride = T;
    uint32_t dmax = 0;
    for (uint32_t t = 0; t < T; ++t) dmax = std::max(dmax, distances[t]);
    ao.maxdist = (float)max_seconds * speed_m_s;  // data.rs:532
    // the launch takes at most kMaxT thresholds: chunk t0 covers t0 .. t0 + T - 1
    // of the rows (every chunk searches to the same walk distance)
    auto chunk = [&](uint32_t t0) {
      AccessOut c = ao;
      c.T = chunk_len(T, t0);
      c.max_mask = 0;
      for (uint32_t t = 0; t < c.T; ++t) {
        c.dist[t] = (float)distances[t0 + t];
        c.beta[t] = betas[t0 + t];
        c.mcw[t] = max_curve_wts[t0 + t];
        if (distances[t0 + t] == dmax) c.max_mask |= 1u << t;
      }
      c.count = ao.count + (size_t)t0 * N;
      c.weighted = ao.weighted ? ao.weighted + (size_t)t0 * N : nullptr;
      if (ao.stats) c.stats = ao.stats + (size_t)t0 * N;
      return c;
    };
    const int dev = current_device();
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, dev);
    DevEdges &de = device_edges(g, dg, speed_m_s);
    OwnedStream stream;
    DevBuf d_node_off(4ull * (N + 1)), d_node_mem(4ull * n_assign), d_mem(16ull * n_assign),
        d_class(4ull * std::max<uint32_t>(n_entries, 1)),
        d_src(4ull * std::max<size_t>(srcs.size(), 1)), d_count(8 * CTN), d_wt(out_div ? 8 : 8 * CTN),
        d_near(8ull * C * N);
    auto up = [&](DevBuf &d, const void *hsrc, size_t bytes) {
      if (bytes) hip_check(hipMemcpyAsync(d.p, hsrc, bytes, hipMemcpyHostToDevice, stream), "copy");
    };
    up(d_node_off, node_off.data(), 4ull * (N + 1));
    up(d_node_mem, node_mem.data(), 4ull * n_assign);
    up(d_mem, mem.data(), 16ull * n_assign);
    if (entry_class) up(d_class, entry_class, 4ull * n_entries);
    up(d_src, srcs.data(), 4ull * srcs.size());
    {
      const std::vector<double> nan_rows((size_t)C * N, std::nan(""));  // NaN bits: no entry yet
      up(d_near, nan_rows.data(), 8ull * C * N);
    }
    hip_check(hipMemsetAsync(d_count.p, 0, 8 * CTN, stream), "memset");
    if (!out_div) hip_check(hipMemsetAsync(d_wt.p, 0, 8 * CTN, stream), "memset");
    ao.node_off = d_node_off.as<uint32_t>();
    ao.node_mem = d_node_mem.as<uint32_t>();
    ao.mem = d_mem.as<uint4>();
    ao.n_shared = n_shared;
    ao.n_groups = n_groups;
    DevBuf d_stats(8ull * std::max<size_t>((size_t)SM * T * N, 1)),
        d_values(4ull * std::max<size_t>(stats_in ? (size_t)stats_in->n_maps * n_entries : 0, 1));
    if (stats_in) {
      up(d_stats, stat_init.data(), 8ull * SM * T * N);
      up(d_values, stats_in->values, 4ull * stats_in->n_maps * n_entries);
      ao.stats = d_stats.as<double>();
      ao.values = d_values.as<float>();
      ao.n_maps = stats_in->n_maps;
      ao.n_entries = n_entries;
    }
    ao.entry_class = d_class.as<int32_t>();
    ao.count = d_count.as<double>();
    ao.weighted = out_div ? nullptr : d_wt.as<double>();
    ao.nearest = d_near.as<unsigned long long>();
    uint32_t *prog = progress_begin(g, N - srcs.size());
    if (!srcs.empty() && !angular)
      for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
        launch_access(g, dg, de, (float)max_seconds, d_src.as<uint32_t>(), (uint32_t)srcs.size(), chunk(t0), stream,
                      t_stats, t0 + kMaxT >= T ? prog : nullptr);
    if (!srcs.empty() && angular) {
      // angular aggregation walks dijkstra_tree_simplest (data.rs:538-548):
      // each source's reached nodes come from the simplest-path reach kernel,
      // then the same aggregation runs over those lists. Sources go in rounds:
      // a round's lists hold `cap` nodes per source, and sources that reach
      // more (or fill the heap) go to the next round with four times the room,
      // up to the whole graph (2 M + 2 heap entries: one push per relaxation).
      const uint32_t hmax = (uint32_t)std::min<uint64_t>(2ull * g.m + 2, 0xfffffff0ull);
      // the tuning's slot capacity (search states per lane) bounds the first round's lists at half as many nodes
      const uint32_t cap0 = default_tuning().cap;
      uint32_t cap = std::min<uint32_t>(N, cap0 ? std::max<uint32_t>(cap0 / 2, 1u) : 4096u);
      uint32_t hcap = (uint32_t)std::min<uint64_t>(hmax, 4ull * cap);
      bool heap = false;  // the bucket ring first; the heap once a bucket overfills
      std::vector<uint32_t> pending(srcs.size());
      for (uint32_t i = 0; i < (uint32_t)pending.size(); ++i) pending[i] = i;
      while (!pending.empty()) {
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "meminfo");
        const size_t per = 8ull * cap;
        const size_t chunk_n = std::max<size_t>(64, std::min<size_t>(pending.size(), free_b / 8 / per));
        DevBuf d_items(4 * chunk_n), d_counts(4 * chunk_n), d_lists(per * chunk_n);
        std::vector<uint32_t> next;
        uint32_t bits = 0;
        for (size_t c0 = 0; c0 < pending.size(); c0 += chunk_n) {
          const uint32_t n = (uint32_t)std::min<size_t>(chunk_n, pending.size() - c0);
          hip_check(hipMemcpyAsync(d_items.p, pending.data() + c0, 4ull * n, hipMemcpyHostToDevice, stream), "copy");
          uint32_t b = 0;
          for (uint32_t f : launch_reach_simplest(g, dg, de, (float)max_seconds, d_src.as<uint32_t>(),
                                                  d_items.as<uint32_t>(), n, cap, hcap, heap, d_lists.as<uint64_t>(),
                                                  d_counts.as<uint32_t>(), stream, t_stats, &b))
            next.push_back(pending[c0 + f]);
          bits |= b;
          for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
            launch_access_lists(chunk(t0), d_src.as<uint32_t>(), d_items.as<uint32_t>(), d_lists.as<uint64_t>(),
                                d_counts.as<uint32_t>(), n, cap, (float)max_seconds, speed_m_s, stream, t_stats,
                                t0 + kMaxT >= T ? prog : nullptr);
        }
        if (next.empty()) break;
        if (cap >= N && heap && hcap >= hmax) fail(CS_EINTERNAL, "angular reach lists: capacity retries exhausted");
        t_stats.retries += next.size();
        if (bits & kReachFailCap) cap = (uint32_t)std::min<uint64_t>(N, 4ull * cap);
        const bool grow_heap = heap && (bits & kReachFailHeap);
        if (bits & kReachFailHeap) heap = true;
        hcap = (uint32_t)std::min<uint64_t>(hmax, std::max<uint64_t>(4ull * cap, grow_heap ? 4ull * hcap : hcap));
        pending.swap(next);
      }
    }
    const uint32_t *ip = result_order(g);
    if (stats_in) {
      std::vector<double> hs((size_t)SM * T * N);
      hip_check(hipMemcpyAsync(hs.data(), d_stats.p, 8 * hs.size(), hipMemcpyDeviceToHost, stream), "copy out");
      hip_check(hipStreamSynchronize(stream), "sync");
      unpack_acc(hs, N, SM, T, stats_in->out, ip);
      t_stats.sources = srcs.size();
      progress_finish(g, N);
      return;
    }
    if (out_div) {
      DevBuf d_div(8ull * kDivMetrics * T * N);
      for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
        launch_diversity(chunk(t0), C, d_div.as<double>() + (size_t)t0 * N, stream);
      std::vector<double> hd((size_t)kDivMetrics * T * N);
      hip_check(hipMemcpyAsync(hd.data(), d_div.p, 8 * hd.size(), hipMemcpyDeviceToHost, stream), "copy out");
      hip_check(hipStreamSynchronize(stream), "sync");
      unpack_acc(hd, N, kDivMetrics, T, out_div, ip);
      t_stats.sources = srcs.size();
      progress_finish(g, N);
      return;
    }
    std::vector<double> hc(CTN), hw(CTN), hn((size_t)C * N);
    hip_check(hipMemcpyAsync(hc.data(), d_count.p, 8 * CTN, hipMemcpyDeviceToHost, stream), "copy out");
    hip_check(hipMemcpyAsync(hw.data(), d_wt.p, 8 * CTN, hipMemcpyDeviceToHost, stream), "copy out");
    hip_check(hipMemcpyAsync(hn.data(), d_near.p, 8ull * C * N, hipMemcpyDeviceToHost, stream), "copy out");
    hip_check(hipStreamSynchronize(stream), "sync");
    unpack_acc(hc, N, C, T, out_count, ip);
    unpack_acc(hw, N, C, T, out_weighted, ip);
    unpack_acc(hn, N, C, 1, out_nearest, ip);
    t_stats.sources = srcs.size();
    progress_finish(g, N);
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_accessibility(cs_graph *h, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                     uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node,
                     const uint32_t *assign_entry, const double *assign_dist, uint32_t n_thresholds,
                     const uint32_t *distances, const float *betas, const float *max_curve_wts,
                     uint32_t max_seconds, float speed_m_s, int angular, double *out_count,
                     double *out_weighted, double *out_nearest) {
  return access_impl(h, n_entries, entry_group, entry_class, n_classes, n_assign, assign_node, assign_entry,
                     assign_dist, n_thresholds, distances, betas, max_curve_wts, max_seconds, speed_m_s, angular,
                     out_count, out_weighted, out_nearest, nullptr);
}

int cs_mixed_uses(cs_graph *h, uint32_t n_entries, const uint32_t *entry_group, const int32_t *entry_class,
                  uint32_t n_classes, uint32_t n_assign, const uint32_t *assign_node, const uint32_t *assign_entry,
                  const double *assign_dist, uint32_t n_thresholds, const uint32_t *distances, const float *betas,
                  const float *max_curve_wts, uint32_t max_seconds, float speed_m_s, int angular, double *out) {
  if (!out) return guarded([] { fail(CS_EINVAL, "out must not be NULL"); });
  return access_impl(h, n_entries, entry_group, entry_class, n_classes, n_assign, assign_node, assign_entry,
                     assign_dist, n_thresholds, distances, betas, max_curve_wts, max_seconds, speed_m_s, angular,
                     nullptr, nullptr, nullptr, out);
}

int cs_data_stats(cs_graph *h, uint32_t n_entries, const uint32_t *entry_group, uint32_t n_assign,
                  const uint32_t *assign_node, const uint32_t *assign_entry, const double *assign_dist,
                  uint32_t n_maps, const float *values, uint32_t n_thresholds, const uint32_t *distances,
                  const float *betas, const float *max_curve_wts, uint32_t max_seconds, float speed_m_s,
                  int angular, double *out) {
  if (!out) return guarded([] { fail(CS_EINVAL, "out must not be NULL"); });
  StatsIn st;
  st.n_maps = n_maps;
  st.values = values;
  st.out = out;
  return access_impl(h, n_entries, entry_group, nullptr, 0, n_assign, assign_node, assign_entry, assign_dist,
                     n_thresholds, distances, betas, max_curve_wts, max_seconds, speed_m_s, angular, nullptr, nullptr,
                     nullptr, nullptr, &st);
}

int cs_tree_shortest(cs_graph *h, uint32_t src, uint32_t max_seconds, float speed, uint8_t *visited,
                     uint8_t *discovered, int32_t *pred, float *short_dist, float *agg_seconds, uint32_t *order,
                     uint32_t *n_order) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    Single s = run_single(h, src, max_seconds, speed, kTieEps, 0, 0);
    const uint32_t N = h->g.n;
    for (uint32_t v = 0; v < N; ++v) {
      const bool r = std::isfinite(s.secs[v]);
      visited[v] = r;
      discovered[v] = r;
      pred[v] = s.pred[v];
      agg_seconds[v] = s.secs[v];
      short_dist[v] = v == src ? 0.0f : (r ? s.secs[v] * speed : INFINITY);
    }
    const uint32_t k = settle_order(s.secs, order);
    if (n_order) *n_order = k;
  });
}

int cs_tree_segment(cs_graph *h, uint32_t src, uint32_t max_seconds, float speed, uint8_t *visited,
                    uint8_t *discovered, int32_t *pred, float *short_dist, float *agg_seconds,
                    int32_t *origin_seg, int32_t *last_seg, uint32_t *order, uint32_t *n_order,
                    uint32_t *edge_order, uint32_t *n_edges) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    Single s = run_single(h, src, max_seconds, speed, kTieEps, 0, 0);
    const Graph &g = h->g;
    const uint32_t N = g.n;
    std::vector<uint32_t> ord(N);
    const uint32_t k = settle_order(s.secs, ord.data());
    for (uint32_t v = 0; v < N; ++v) {
      const bool r = std::isfinite(s.secs[v]);
      visited[v] = r;
      discovered[v] = r;
      pred[v] = s.pred[v];
      agg_seconds[v] = s.secs[v];
      short_dist[v] = v == src ? 0.0f : (r ? s.secs[v] * speed : INFINITY);
      last_seg[v] = s.last[v];
      origin_seg[v] = -1;
    }
    // origin segment: the first edge of each tree path, propagated in settle order
    for (uint32_t i = 0; i < k; ++i) {
      const uint32_t v = ord[i];
      if (v == src || s.pred[v] < 0) continue;
      origin_seg[v] = (uint32_t)s.pred[v] == src ? s.last[v] : origin_seg[s.pred[v]];
    }
    // visited edges: in-edges of each settled node whose source had not been
    // settled yet, in settle order and petgraph order (centrality.rs:1553-1577)
    uint32_t ne = 0;
    std::vector<uint32_t> pos(N, UINT32_MAX);
    for (uint32_t i = 0; i < k; ++i) pos[ord[i]] = i;
    for (uint32_t i = 0; i < k; ++i) {
      const uint32_t v = ord[i];
      for (uint32_t j = g.rev_off[v]; j < g.rev_off[v + 1]; ++j) {
        const uint32_t e = g.rev_eid[j];
        const uint32_t nb = g.esrc[e];
        if (nb != v && pos[nb] != UINT32_MAX && pos[nb] < i) continue;
        if (edge_order) edge_order[ne] = e;
        ++ne;
      }
    }
    if (order) std::copy(ord.begin(), ord.begin() + k, order);
    if (n_order) *n_order = k;
    if (n_edges) *n_edges = ne;
  });
}

int cs_brandes_shortest(cs_graph *h, uint32_t src, uint32_t max_seconds, float speed, float tolerance,
                        float *secs, float *route, double *sigma, uint32_t *n_preds, uint32_t *preds,
                        uint32_t max_preds) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    if (!(tolerance >= 0.0f && tolerance <= 1.0f))
      fail(CS_EINVAL, "Tolerance must be between 0 and 100 (percent), got " + std::to_string(tolerance * 100.0f));
    Single s = run_single(h, src, max_seconds, speed, tolerance, 1, max_preds);
    const uint32_t N = h->g.n;
    for (uint32_t v = 0; v < N; ++v) {
      secs[v] = s.secs[v];
      route[v] = std::isfinite(s.secs[v]) ? s.secs[v] * speed : INFINITY;
      sigma[v] = s.sigma[v];
      n_preds[v] = s.npreds[v];
    }
    if (max_preds) std::copy(s.preds.begin(), s.preds.end(), preds);
  });
}

int cs_tree_simplest(cs_graph *h, uint32_t src, uint32_t max_seconds, float speed, uint8_t *visited,
                     uint8_t *discovered, int32_t *pred, float *simpl_dist, float *agg_seconds, uint32_t *order,
                     uint32_t *n_order) {
  return guarded([&] {
    if (!h) fail(CS_EINVAL, "graph must not be NULL");
    Graph &g = h->g;
    if (!g.is_dual)
      fail(CS_EINVAL,
           "dijkstra_tree_simplest requires a dual graph for angular analysis. Convert the graph with "
           "cityseer.tools.graphs.nx_to_dual(...) before ingesting it into NetworkStructure.");
    if (src >= g.n)
      fail(CS_EINVAL, "src_idx " + std::to_string(src) + " out of range for network with node_bound " +
                          std::to_string(g.n));
    if (!std::isfinite(speed) || !(speed > 0.0f)) fail(CS_EINVAL, "speed_m_s must be finite and positive");
    if (!g.slots_ok) fail(CS_EINVAL, g.slots_err);
    const int dev = current_device();
    const uint32_t N = g.n;
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, dev);
    DevEdges &de = device_edges(g, dg, speed);
    OwnedStream stream;
    std::vector<float> hs(N, INFINITY), hsec(N, INFINITY);
    std::vector<int32_t> hp(N, -1);
    std::vector<uint8_t> hv(N, 0);
    std::vector<uint32_t> ho(N);
    uint32_t hn = 0;
    DevBuf d_s(4ull * N), d_sec(4ull * N), d_p(4ull * N), d_v(N), d_o(4ull * N), d_n(4);
    hip_check(hipMemcpyAsync(d_s.p, hs.data(), 4ull * N, hipMemcpyHostToDevice, stream), "copy");
    hip_check(hipMemcpyAsync(d_sec.p, hsec.data(), 4ull * N, hipMemcpyHostToDevice, stream), "copy");
    hip_check(hipMemcpyAsync(d_p.p, hp.data(), 4ull * N, hipMemcpyHostToDevice, stream), "copy");
    hip_check(hipMemsetAsync(d_v.p, 0, N, stream), "memset");
    launch_tree_simplest(g, dg, de, src, (float)max_seconds, d_s.as<float>(), d_sec.as<float>(), d_p.as<int32_t>(),
                         d_v.as<uint8_t>(), d_o.as<uint32_t>(), d_n.as<uint32_t>(), stream);
    hip_check(hipMemcpyAsync(hs.data(), d_s.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(hsec.data(), d_sec.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(hp.data(), d_p.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(hv.data(), d_v.p, N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(ho.data(), d_o.p, 4ull * N, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipMemcpyAsync(&hn, d_n.p, 4, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "sync");
    for (uint32_t v = 0; v < N; ++v) {
      visited[v] = hv[v];
      discovered[v] = hv[v];
      pred[v] = hp[v];
      simpl_dist[v] = hs[v];
      agg_seconds[v] = hsec[v];
    }
    if (order) std::copy(ho.begin(), ho.begin() + hn, order);
    if (n_order) *n_order = hn;
  });
}

}  // extern "C"

void cs_batch::build() {
  if (sb) return;
  if (!stream) hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
  acc_words = 0;
  for (auto &j : jobs) {
    j.acc_off = acc_words;
    acc_words += (size_t)j.h->g.n * 7 * NT;
  }
  hip_check(hipMalloc(&d_acc, 8 * std::max<size_t>(acc_words, 1)), "hipMalloc batch acc");
  hip_check(hipMalloc(&d_reach, 32 * std::max<size_t>(jobs.size(), 1)), "hipMalloc batch reach");
  std::vector<JobSpec> specs;
  for (size_t i = 0; i < jobs.size(); ++i) {
    JobHost &j = jobs[i];
    Graph &g = j.h->g;
    const uint32_t n = (uint32_t)j.ss.src.size();
    const uint32_t *d_src = upload(j.ss.src.data(), n);
    const float *d_wt = upload(j.ss.wt.data(), n);
    const float *d_cyc = upload(j.ss.cyc.data(), n);
    const uint8_t *d_el = upload(j.elig.data(), j.elig.size());
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph &dg = device_graph(g, device);
    DevEdges &de = device_edges(g, dg, j.spec.speed);
    specs.push_back(JobSpec{&g, &dg, &de, j.spec, d_src, d_wt, d_cyc, n, d_el, d_acc + j.acc_off,
                            d_reach + 8 * i, tune.delta_scale});
  }
  std::vector<ZeroRegion> zero{{reinterpret_cast<uint64_t *>(d_acc), acc_words},
                               {reinterpret_cast<uint64_t *>(d_reach), 4 * jobs.size()}};
  sb = batch_new(std::move(specs), std::move(zero), NT, device, tune);
}

namespace {
Tuning from_c(const cs_tuning *t) {
  if (!t) fail(CS_EINVAL, "tuning must not be NULL");
  if (t->block_threads != 0 && (t->block_threads % 64 != 0 || t->block_threads > 256))
    fail(CS_EINVAL, "block_threads must be 64, 128, 192 or 256");
  if (!(t->delta_scale > 0.0f && t->delta_scale <= 1.0f)) fail(CS_EINVAL, "delta_scale must be in (0, 1]");
  Tuning r;
  r.block = t->block_threads ? t->block_threads : 256;
  r.waves_per_simd = t->waves_per_simd;
  r.delta_scale = t->delta_scale;
  r.cap = t->slot_capacity;
  r.xcd_split = t->xcd_split ? 1u : 0u;
  if (t->simplest_lanes > 64) fail(CS_EINVAL, "simplest_lanes must be at most 64");
  r.ang_lanes = t->simplest_lanes;
  return r;
}
cs_tuning to_c(const Tuning &t) {
  cs_tuning r;
  r.block_threads = t.block;
  r.waves_per_simd = t.waves_per_simd;
  r.delta_scale = t.delta_scale;
  r.slot_capacity = t.cap;
  r.xcd_split = t.xcd_split;
  r.simplest_lanes = t.ang_lanes;
  return r;
}
}  // namespace

extern "C" {

int cs_set_tuning(const cs_tuning *t) {
  return guarded([&] { set_default_tuning(from_c(t)); });
}

int cs_release_workspaces(void) {
  return guarded([] {
    release_search_workspaces();
    release_angular_workspaces();
  });
}

int cs_get_tuning(cs_tuning *t) {
  return guarded([&] {
    if (!t) fail(CS_EINVAL, "tuning must not be NULL");
    *t = to_c(default_tuning());
  });
}

int cs_batch_create(cs_batch **out) {
  return guarded([&] {
    if (!out) fail(CS_EINVAL, "out must not be NULL");
    auto b = std::make_unique<cs_batch>();
    b->device = current_device();
    b->tune = default_tuning();
    *out = b.release();
  });
}

void cs_batch_destroy(cs_batch *b) {
  delete b;
  (void)hipGetLastError();
}

int cs_batch_add_shortest(cs_batch *b, cs_graph *h, const cs_params *p, double *out, uint32_t *reach_totals) {
  int idx = -1;
  int rc = guarded([&] {
    if (!b || !h || !out) fail(CS_EINVAL, "batch, graph and out must not be NULL");
    Graph &g = h->g;
    check_params(g, p);
    run_shortest_checks(g, p);
    SourceSet ss = source_set(g, p);
    to_device_sources(g, ss);
    const std::vector<uint8_t> elig = to_device_u8(g, p->eligible);
    // more than kMaxT thresholds: one job per chunk, each searching to the
    // request's largest cutoff (make_spec) and filling its own rows of `out`
    const uint32_t T = p->n_thresholds;
    std::vector<cs_batch::JobHost> chunks;
    for (uint32_t t0 = 0; t0 < T; t0 += kMaxT) {
      cs_batch::JobHost j;
      j.h = h;
      j.spec = make_spec(p, t0, chunk_len(T, t0));
      j.ss = ss;
      j.elig = elig;
      j.out = out;
      j.reach = reach_totals;
      j.betw = p->compute_betweenness != 0;
      j.betw_scale = p->betw_scale;
      j.out_T = T;
      j.t0 = t0;
      chunks.push_back(std::move(j));
    }
    b->drop_device();
    for (auto &j : chunks) {
      b->NT = std::max(b->NT, j.spec.T);
      b->jobs.push_back(std::move(j));
    }
    idx = b->n_requests++;
  });
  return rc == CS_OK ? idx : -rc;
}

int cs_batch_set_tuning(cs_batch *b, const cs_tuning *t) {
  return guarded([&] {
    if (!b) fail(CS_EINVAL, "batch must not be NULL");
    b->tune = from_c(t);
    if (b->sb) batch_set_tuning(*b->sb, b->tune);
  });
}

int cs_batch_autotune(cs_batch *b, double sample_fraction, cs_tuning *chosen) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!b) fail(CS_EINVAL, "batch must not be NULL");
    if (!(sample_fraction > 0.0 && sample_fraction <= 1.0)) fail(CS_EINVAL, "sample_fraction must be in (0, 1]");
    b->build();
    b->tune = batch_autotune(*b->sb, sample_fraction, b->stream);
    b->have_results = false;
    if (chosen) *chosen = to_c(b->tune);
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_batch_run(cs_batch *b) {
  const auto t0 = std::chrono::steady_clock::now();
  t_stats = cs_stats{};
  int rc = guarded([&] {
    if (!b) fail(CS_EINVAL, "batch must not be NULL");
    b->build();
    batch_run(*b->sb, b->stream, t_stats);
    uint64_t n = 0;
    for (auto &j : b->jobs) n += j.ss.src.size();
    t_stats.sources = n;
    b->have_results = true;
  });
  t_stats.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int cs_batch_results(cs_batch *b) {
  return guarded([&] {
    if (!b) fail(CS_EINVAL, "batch must not be NULL");
    if (!b->have_results) fail(CS_EINVAL, "the batch has no results: call cs_batch_run first");
    std::vector<double> hacc(b->acc_words);
    std::vector<uint32_t> hreach(8 * b->jobs.size());
    if (b->acc_words)
      hip_check(hipMemcpy(hacc.data(), b->d_acc, 8 * b->acc_words, hipMemcpyDeviceToHost), "copy out");
    if (!hreach.empty())
      hip_check(hipMemcpy(hreach.data(), b->d_reach, 4 * hreach.size(), hipMemcpyDeviceToHost), "copy out");
    for (size_t i = 0; i < b->jobs.size(); ++i) {
      const cs_batch::JobHost &j = b->jobs[i];
      const uint32_t N = j.h->g.n, T = j.spec.T;
      unpack_acc_strided(hacc.data() + j.acc_off, N, 7, T, b->NT, j.out, result_order(j.h->g), j.out_T, j.t0);
      if (j.betw) scale_rows(j.out, N, j.out_T, 5, 6, j.betw_scale, j.t0, T);
      if (j.reach)
        for (uint32_t t = 0; t < T; ++t) j.reach[j.t0 + t] = hreach[8 * i + t];
    }
  });
}

}  // extern "C"
