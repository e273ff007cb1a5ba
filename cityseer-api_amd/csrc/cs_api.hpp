// cs_api.hpp -- host helpers shared by the C ABI translation units
// (cs_host.cpp: graphs, plain calls and batches; cs_serve.cpp: the request
// server): parameter checks, source preparation and result unpacking.
#pragma once

#include <algorithm>
#include <functional>
#include <string>
#include <vector>

#include "cs_internal.hpp"

struct cs_graph {
  cs::Graph g;
};

namespace cs {

// Current HIP device; raises CS_EDEVICE without a GPU.
int current_device();

// Runs f, turning exceptions into a status code and the thread's last error.
int guarded_call(const std::function<void()> &f);

// Argument checks shared by the centrality entry points (the reference's
// ValueError messages, centrality.rs:895-959, common.rs:98-270).
void check_params(const Graph &g, const cs_params *p);
void run_shortest_checks(const Graph &g, const cs_params *p);
void check_betweenness_graph(const Graph &g);
void check_simplest_graph(const Graph &g);
// The kernels take at most kMaxT thresholds per launch; a call with more runs
// one launch per chunk of kMaxT, each writing its own threshold rows.
// make_spec(p, t0, n) holds thresholds t0 .. t0 + n - 1 (n = 0: the rest,
// at most kMaxT) and the call's largest cutoff.
CallSpec make_spec(const cs_params *p, uint32_t t0 = 0, uint32_t n = 0);
inline uint32_t chunk_len(uint32_t T, uint32_t t0) { return std::min<uint32_t>(kMaxT, T - t0); }

// Sources kept after NaN-weight filtering, with per-source weights and the
// cycles weight wt / node_weight (centrality.rs:1711-1730).
struct SourceSet {
  std::vector<uint32_t> src;
  std::vector<float> wt, cyc;
};
SourceSet source_set(const Graph &g, const cs_params *p);

// Device node order (Graph::perm). The shortest and segment searches run on
// device node ids: sources are translated and sorted ascending (waves that run
// side by side then search overlapping neighbourhoods), per-node inputs are
// gathered into device order and results scattered back on unpacking. The
// angular search keeps the original order (dual graphs are never renumbered).
void to_device_sources(const Graph &g, SourceSet &ss);
void to_device_ids(const Graph &g, std::vector<uint32_t> &ids);
std::vector<uint8_t> to_device_u8(const Graph &g, const uint8_t *in);
// iperm to pass to unpack_acc_strided: null when the order is the identity
const uint32_t *result_order(const Graph &g);

// A non-blocking stream owned by one call: destroyed on every exit path,
// after the call's device buffers (declared later) have been freed.
struct OwnedStream {
  hipStream_t s = nullptr;
  OwnedStream() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream"); }
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
  OwnedStream(const OwnedStream &) = delete;
  OwnedStream &operator=(const OwnedStream &) = delete;
  operator hipStream_t() const { return s; }
};

// acc [M][stride][N] (device rows, node d at column d) -> rows t0 .. t0 + T - 1
// of out [M][out_T][N] (host, node iperm[d] at column iperm[d]; iperm null =
// identity; out_T 0 = T)
void unpack_acc_strided(const double *acc, uint32_t N, uint32_t M, uint32_t T, uint32_t stride, double *out,
                        const uint32_t *iperm, uint32_t out_T = 0, uint32_t t0 = 0);
// scales rows m0..m1 of out [M][T][N], thresholds t0 .. t0 + n_t - 1 (n_t 0 = the rest)
void scale_rows(double *out, uint32_t N, uint32_t T, uint32_t m0, uint32_t m1, double s, uint32_t t0 = 0,
                uint32_t n_t = 0);

}  // namespace cs
