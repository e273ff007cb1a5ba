// cs_noalloc.hip -- the allocation-free form of the device-resident
// shortest-path entry point. Its working memory is one static device array per
// device (`scratch`); the host takes that array's device address once per
// device with hipGetSymbolAddress and hands that address, never the array's
// name, to kernels.
#include <chrono>
#include <cmath>
#include <mutex>
#include <string>

#include "cs_api.hpp"

namespace {

constexpr size_t kScratchBytes = CS_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

}  // namespace

// The working memory of cs_centrality_shortest_device_noalloc: each device
// holds its own copy at its own address.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};
// One call at a time per device works in its array (a call returns only after
// its stream has finished with it).
std::mutex g_device_mu[kMaxDevices];

// Device address of this device's `scratch`, taken on the first call on the
// device (with it current) and kept.
char *scratch_on(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      cs::fail(CS_EINTERNAL, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char *>(ptr);
  }
  return g_scratch_ptr[device];
}

}  // namespace

using namespace cs;

int cs_centrality_shortest_device_noalloc(cs_graph *h, const cs_params *p, const float *d_cycles_wt, double *d_acc,
                                          uint32_t *d_reach, void *stream_ptr) {
  const auto t0 = std::chrono::steady_clock::now();
  cs_stats &st = thread_stats();
  st = cs_stats{};
  const int rc = guarded_call([&] {
    if (!h || !p || !d_acc) fail(CS_EINVAL, "graph, params and d_acc must not be NULL");
    Graph &g = h->g;
    if (p->n_thresholds == 0) fail(CS_EINVAL, "at least one distance threshold is required");
    if (!p->compute_closeness && !p->compute_betweenness)
      fail(CS_EINVAL, "Either or both closeness and betweenness flags is required, but both parameters are False.");
    if (!p->sources || !p->source_wt || !d_cycles_wt || !p->eligible)
      fail(CS_EINVAL, "device sources, weights, cycle weights and eligibility must not be NULL");
    const int dev = current_device();
    const uint32_t T = p->n_thresholds;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    if (dev >= kMaxDevices) fail(CS_EDEVICE, "device index out of range");
    char *ws = scratch_on(dev);
    std::lock_guard<std::mutex> dlk(g_device_mu[dev]);
    std::lock_guard<std::mutex> lk(g.mu);
    DevGraph *dg = find_device_graph(g, dev);
    DevEdges *de = dg ? find_device_edges(*dg, p->speed_m_s) : nullptr;
    if (!de)
      fail(CS_EINVAL, "the graph is not resident on device " + std::to_string(dev) + " at speed " +
                          std::to_string(p->speed_m_s) + ": call cs_graph_upload first");
    run_shortest_checks(g, p);
    for (uint32_t t0 = 0; t0 < T; t0 += kMaxT)
      launch_shortest_in(g, *dg, *de, make_spec(p, t0, chunk_len(T, t0)), p->sources, p->source_wt, d_cycles_wt,
                         p->n_sources, p->eligible, d_acc, d_reach, T, t0, stream, st, ws, kScratchBytes);
    st.sources = p->n_sources;
  });
  st.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
