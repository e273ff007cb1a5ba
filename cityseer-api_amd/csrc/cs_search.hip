// cs_search.hip -- distance-bounded per-source search on gfx950.
//
// One wavefront (64 lanes) owns one source at a time and pulls the next source
// from a global counter when it finishes (persistent waves, dynamic balance).
// The search is a Delta-bucketed Dijkstra over the reverse graph, the
// direction the reference searches (rust/src/centrality.rs:1389-1392):
//
//  * Delta is below the smallest edge cost by a margin, so a node settled in
//    bucket k only relaxes into buckets > k, and every node of bucket k is
//    final when bucket k is opened. All nodes of a bucket settle in parallel.
//  * The margin also guarantees that a candidate offered by a node of the same
//    bucket can never enter the tie band of the receiver: the reference's
//    order-dependent multi-predecessor fold (centrality.rs:1413-1437) can
//    therefore be replayed per node from candidates of earlier buckets only,
//    in the reference's settle order (ascending seconds, then node index),
//    reproducing predecessor sets and sigma bit-exactly.
//  * Per-wave node state is an epoch-tagged 64-bit key (epoch<<32 | ~f32 bits
//    of seconds) updated with a 64-bit atomic max: newer epochs and smaller
//    seconds win, zero-filled memory is stale, and no O(N) state is cleared
//    per source. Settled nodes get compact slots in settle order.
//  * Closeness terms are scattered to targets with f64 atomics
//    (centrality.rs:1755-1778); betweenness back-propagates bucket by bucket
//    in reverse (centrality.rs:823-873).
#include "cs_internal.hpp"

#include <cmath>
#include <cstring>

namespace cs {

namespace {

constexpr uint32_t kFailCap = 1u, kFailBucket = 2u, kFailRing = 4u, kFailInternal = 8u;

enum Mode { kBrandes = 0, kSegment = 1, kTree = 2, kBrandesExport = 3 };

struct SearchArgs {
  uint32_t N;
  const uint32_t *rev_off;
  const uint2 *rev_adj;
  const uint32_t *fwd_off;
  const uint2 *fwd_adj;
  const float2 *rev_tw;
  const float *fwd_twlen;
  const uint32_t *fwd_eid;
  const uint8_t *eligible;
  uint32_t T;
  float maxsec, speed, inv_delta, tolerance;
  float dist[kMaxT];
  float beta[kMaxT];
  uint32_t nbk, bcap, cap, maxb;
  uint32_t n_sources;
  const uint32_t *sources;
  const float *src_wt;
  const float *src_cyc;
  uint32_t flags;  // 1 closeness, 2 betweenness, 4 track reach
  double *acc;     // [N][M][T]
  uint32_t *reach;
  unsigned long long *stats;  // reached, edges, within
  char *ws;
  size_t ws_stride;
  size_t o_keyw, o_slotw, o_node, o_secs, o_a8, o_b8, o_acc, o_bkt, o_bcnt, o_bnd;
  uint32_t *epochs;
  uint32_t *next_source;
  uint32_t *fail_list;
  uint32_t *n_fail;
  uint32_t *fail_flags;
  TreeOut tout;
};

// Wave-private workspace: only the owning wave reads or writes it, so
// workgroup scope (coherent in the CU's L1 and the XCD's L2) is sufficient.
__device__ __forceinline__ uint64_t ld64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ float ldf(const float *p) {
  return __uint_as_float(ld32(reinterpret_cast<const uint32_t *>(p)));
}
__device__ __forceinline__ double ldd(const double *p) {
  return __longlong_as_double((long long)ld64(reinterpret_cast<const uint64_t *>(p)));
}
__device__ __forceinline__ void st32(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st64(uint64_t *p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void std_(double *p, double v) {
  st64(reinterpret_cast<uint64_t *>(p), (uint64_t)__double_as_longlong(v));
}
__device__ __forceinline__ uint64_t wmax64(uint64_t *p, uint64_t v) {
  return __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t wadd32(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void wadd_f64(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Results shared by all waves (accumulators): device scope.
__device__ __forceinline__ void gadd_f64(double *p, double v) { unsafeAtomicAdd(p, v); }
// Orders this wave's earlier global writes before its later reads (all lanes).
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t pack_key(uint32_t epoch, float secs) {
  return ((uint64_t)epoch << 32) | (uint32_t)(~__float_as_uint(secs));
}
__device__ __forceinline__ uint64_t pack_key_bits(uint32_t epoch, uint32_t secs_bits) {
  return ((uint64_t)epoch << 32) | (uint32_t)(~secs_bits);
}
__device__ __forceinline__ bool key_current(uint64_t k, uint32_t epoch) { return (uint32_t)(k >> 32) == epoch; }
__device__ __forceinline__ float key_secs(uint64_t k) { return __uint_as_float(~(uint32_t)k); }
__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint32_t bucket_of(float s, float inv_delta) {
  return (uint32_t)(s * inv_delta);
}

// Reference settle order: ascending seconds (route is monotone in seconds),
// then node index (deterministic stand-in for BinaryHeap's internal order).
__device__ __forceinline__ bool key_less(float sa, uint32_t a, float sb, uint32_t b) {
  return sa < sb || (sa == sb && a < b);
}

struct WaveWs {
  uint64_t *keyw;
  uint32_t *slotw;
  uint32_t *node;
  float *secs;
  uint64_t *a8;
  uint64_t *b8;
  double *acc;
  uint2 *bkt;
  uint32_t *bcnt;
  uint32_t *bnd;
};

__device__ __forceinline__ WaveWs wave_ws(const SearchArgs &a, uint32_t wid) {
  char *base = a.ws + (size_t)wid * a.ws_stride;
  WaveWs w;
  w.keyw = reinterpret_cast<uint64_t *>(base + a.o_keyw);
  w.slotw = reinterpret_cast<uint32_t *>(base + a.o_slotw);
  w.node = reinterpret_cast<uint32_t *>(base + a.o_node);
  w.secs = reinterpret_cast<float *>(base + a.o_secs);
  w.a8 = reinterpret_cast<uint64_t *>(base + a.o_a8);
  w.b8 = reinterpret_cast<uint64_t *>(base + a.o_b8);
  w.acc = reinterpret_cast<double *>(base + a.o_acc);
  w.bkt = reinterpret_cast<uint2 *>(base + a.o_bkt);
  w.bcnt = reinterpret_cast<uint32_t *>(base + a.o_bcnt);
  w.bnd = reinterpret_cast<uint32_t *>(base + a.o_bnd);
  return w;
}

// Multi-predecessor fold of one candidate, exactly as the reference's Phase 1
// relaxation of node nb from settled node u (centrality.rs:1413-1437).
struct Fold {
  float best;
  double sigma;
  uint64_t pmask;
};

__device__ __forceinline__ void fold_apply(Fold &f, float cs, double sigma_u, uint32_t j,
                                           bool u_in_preds) {
  const float one_m = 1.0f - kTieEps;
  const float one_p = 1.0f + kTieEps;
  if (cs < f.best) {
    if (cs < f.best * one_m) {
      f.pmask = 1ull << j;
      f.sigma = sigma_u;
    } else {
      f.pmask |= 1ull << j;
      f.sigma += sigma_u;
    }
    f.best = cs;
  } else if (cs <= f.best * one_p && !u_in_preds) {
    f.pmask |= 1ull << j;
    f.sigma += sigma_u;
  }
}

// A candidate offering more than final * (1+1e-4)^(deg+3) can neither stay a
// predecessor nor change which candidates do: every later reset that must
// remove it also removes anything it influenced (see DESIGN notes on the
// fold). Only "near" candidates need the ordered fold; usually there is one.
__device__ __forceinline__ double far_factor(uint32_t deg) {
  double f = 1.0;
  for (uint32_t i = 0; i < deg + 3; ++i) f *= 1.00010002;
  return f;
}

struct SrcCtx {
  uint32_t src;
  uint32_t tag;
  uint32_t k;      // current absolute bucket
  uint32_t nslot;  // settled nodes so far
  uint32_t nb;     // buckets opened
  uint32_t fail;
};

// ---------------------------------------------------------------------------
// Brandes fold for node v (not the source) over its out-edges v->u: candidates
// are nodes u settled in an earlier bucket with cs = secs_u + es <= maxsec.
// Also counts canonical out-edges to nodes settled before v (circuit rank).
// Returns false on an invariant failure.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool brandes_fold(const SearchArgs &a, const WaveWs &w, uint32_t tag, uint32_t v,
                                             float secs_v, uint32_t kb, bool cycles, Fold &f,
                                             uint32_t &cnt_e) {
  const uint32_t r0 = a.fwd_off[v], r1 = a.fwd_off[v + 1];
  const double thr_far = (double)secs_v * far_factor(r1 - r0);
  uint32_t n_near = 0, near_u = 0, near_j = 0;
  float near_cs = INFINITY;
  for (uint32_t j = r0; j < r1; ++j) {
    const uint2 e = a.fwd_adj[j];
    const uint32_t u = e.x & kNodeMask;
    if (u == v) continue;
    const uint64_t ku = ld64(&w.keyw[u]);
    if (!key_current(ku, tag)) continue;
    const float su = key_secs(ku);
    if (cycles && (e.x & kCanonBit) && key_less(su, u, secs_v, v)) ++cnt_e;
    if (bucket_of(su, a.inv_delta) >= kb) continue;
    const float cs = su + __uint_as_float(e.y);
    if (cs > a.maxsec) continue;
    if ((double)cs <= thr_far) {
      if (n_near == 0) {
        near_u = u;
        near_j = j - r0;
        near_cs = cs;
      }
      ++n_near;
    }
  }
  f.best = INFINITY;
  f.sigma = 0.0;
  f.pmask = 0;
  if (n_near == 0) return false;
  if (n_near == 1) {
    f.best = near_cs;
    f.pmask = 1ull << near_j;
    f.sigma = ldd(reinterpret_cast<const double *>(&w.a8[ld32(&w.slotw[near_u])]));
    return f.best == secs_v;
  }
  // Ordered fold over the near candidates: selection passes in settle order
  // (secs_u, u), ties of one u (parallel edges) in row order.
  bool have_last = false;
  float lsu = 0.0f;
  uint32_t lu = 0, lj = 0;
  for (uint32_t step = 0; step < n_near; ++step) {
    bool found = false;
    float psu = INFINITY, pcs = 0.0f;
    uint32_t pu = 0, pj = 0;
    for (uint32_t j = r0; j < r1; ++j) {
      const uint2 e = a.fwd_adj[j];
      const uint32_t u = e.x & kNodeMask;
      if (u == v) continue;
      const uint64_t ku = ld64(&w.keyw[u]);
      if (!key_current(ku, tag)) continue;
      const float su = key_secs(ku);
      if (bucket_of(su, a.inv_delta) >= kb) continue;
      const float cs = su + __uint_as_float(e.y);
      if (cs > a.maxsec || (double)cs > thr_far) continue;
      const uint32_t jr = j - r0;
      if (have_last && !(key_less(lsu, lu, su, u) || (su == lsu && u == lu && jr > lj))) continue;
      if (!found || key_less(su, u, psu, pu) || (su == psu && u == pu && jr < pj)) {
        found = true;
        psu = su;
        pu = u;
        pj = jr;
        pcs = cs;
      }
    }
    if (!found) return false;
    bool contains = false;
    for (uint64_t m = f.pmask; m; m &= m - 1) {
      const uint32_t jb = __builtin_ctzll(m);
      if ((a.fwd_adj[r0 + jb].x & kNodeMask) == pu) contains = true;
    }
    const double sg = ldd(reinterpret_cast<const double *>(&w.a8[ld32(&w.slotw[pu])]));
    fold_apply(f, pcs, sg, pj, contains);
    have_last = true;
    lsu = psu;
    lu = pu;
    lj = pj;
  }
  return f.best == secs_v;
}

// Single-predecessor fold (dijkstra_tree_shortest / dijkstra_tree_segment,
// centrality.rs:1186-1195,1589-1607): strict < over candidates in settle
// order keeps the first minimum, i.e. the lexicographic min (cs, secs_u, u, j).
__device__ __forceinline__ bool tree_fold(const SearchArgs &a, const WaveWs &w, uint32_t tag, uint32_t v,
                                          float secs_v, uint32_t kb, uint32_t &tree_u, uint32_t &tree_j) {
  const uint32_t r0 = a.fwd_off[v], r1 = a.fwd_off[v + 1];
  float bcs = INFINITY, bsu = INFINITY;
  uint32_t bu = 0xffffffffu, bj = 0xffffffffu;
  for (uint32_t j = r0; j < r1; ++j) {
    const uint2 e = a.fwd_adj[j];
    const uint32_t u = e.x & kNodeMask;
    if (u == v) continue;
    const uint64_t ku = ld64(&w.keyw[u]);
    if (!key_current(ku, tag)) continue;
    const float su = key_secs(ku);
    if (bucket_of(su, a.inv_delta) >= kb) continue;
    const float cs = su + __uint_as_float(e.y);
    if (cs > a.maxsec) continue;
    const uint32_t jr = j - r0;
    if (cs < bcs || (cs == bcs && (su < bsu || (su == bsu && (u < bu || (u == bu && jr < bj)))))) {
      bcs = cs;
      bsu = su;
      bu = u;
      bj = jr;
    }
  }
  tree_u = bu;
  tree_j = bj;
  return bj != 0xffffffffu && bcs == secs_v;
}

// ---------------------------------------------------------------------------
// Search one source: bucket by bucket, each entry of the bucket is validated
// (its key still equals the pushed value), given a slot, folded and relaxed in
// one pass. NT = number of thresholds (compile time).
// Returns false if the source must be retried with a larger workspace.
// ---------------------------------------------------------------------------
template <int MODE, int NT>
__device__ bool search_source(const SearchArgs &a, const WaveWs &w, SrcCtx &c, uint32_t (&nband)[NT],
                              uint32_t (&eband)[NT], unsigned long long &st_edges) {
  const uint32_t lane = lane_id();
  const float speed = a.speed;
  const bool cycles = (MODE == kBrandes) && (a.flags & 1u);
  if (lane == 0) {
    st64(&w.keyw[c.src], pack_key(c.tag, 0.0f));
    st64(reinterpret_cast<uint64_t *>(&w.bkt[0]), (uint64_t)c.src);
    st32(&w.bcnt[0], 1u);
  }
  wave_fence();
  uint32_t pending = 1;
  c.k = 0;
  c.nslot = 0;
  c.nb = 0;
  c.fail = 0;
  while (pending > 0) {
    // every bucket index is below maxsec / delta: a larger k means a broken count
    if (c.k > a.maxb + a.nbk) {
      c.fail |= kFailInternal;
      break;
    }
    const uint32_t ring = c.k % a.nbk;
    const uint32_t cnt = uniform(ld32(&w.bcnt[ring]));
    if (cnt == 0) {
      ++c.k;
      continue;
    }
    if (cnt > a.bcap) c.fail |= kFailBucket;
    if (c.nb >= a.maxb) c.fail |= kFailInternal;
    if (c.fail) break;
    if (lane == 0) {
      st32(&w.bcnt[ring], 0u);
      st32(&w.bnd[c.nb], c.nslot);
    }
    ++c.nb;
    uint32_t pushes = 0, lfail = 0;
    for (uint32_t base = 0; base < cnt; base += kWave) {
      const uint32_t i = base + lane;
      bool valid = false;
      uint32_t v = 0, csb = 0;
      if (i < cnt) {
        const uint64_t ent = ld64(reinterpret_cast<const uint64_t *>(&w.bkt[ring * a.bcap + i]));
        v = (uint32_t)ent;
        csb = (uint32_t)(ent >> 32);
        valid = ld64(&w.keyw[v]) == pack_key_bits(c.tag, csb);
      }
      const uint64_t m = __ballot(valid);
      const uint32_t cntv = (uint32_t)__popcll(m);
      if (c.nslot + cntv > a.cap) {
        c.fail |= kFailCap;
        break;
      }
      const uint32_t s = c.nslot + rank_below(m);
      c.nslot += cntv;
      if (!valid) continue;
      const float secs_v = __uint_as_float(csb);
      const float route_v = secs_v * speed;
      st32(&w.slotw[v], s);
      st32(&w.node[s], v);
      st32(reinterpret_cast<uint32_t *>(&w.secs[s]), csb);
      uint32_t cnt_e = 0;
      if constexpr (MODE == kBrandes || MODE == kBrandesExport) {
        Fold f;
        if (v == c.src) {
          f.sigma = 1.0;
          f.pmask = 0;
        } else if (!brandes_fold(a, w, c.tag, v, secs_v, c.k, cycles, f, cnt_e)) {
          lfail |= kFailInternal;
        }
        st64(&w.a8[s], (uint64_t)__double_as_longlong(f.sigma));
        st64(&w.b8[s], f.pmask);
        if constexpr (MODE == kBrandes) {
#pragma unroll
          for (int t = 0; t < 2 * NT; ++t) std_(&w.acc[(size_t)s * 2 * NT + t], 0.0);
        }
      } else {
        uint32_t tu = 0xffffffffu, tj = 0xffffffffu, pslot = 0xffffffffu;
        float olen = 0.0f, llen = 0.0f;
        if (v != c.src) {
          if (!tree_fold(a, w, c.tag, v, secs_v, c.k, tu, tj)) {
            lfail |= kFailInternal;
          } else {
            pslot = ld32(&w.slotw[tu]);
            llen = a.fwd_twlen[a.fwd_off[v] + tj];
            olen = (tu == c.src) ? llen : __uint_as_float((uint32_t)(ld64(&w.b8[pslot]) >> 32));
          }
        }
        // a8 = {pslot | last-edge row position << 32}; b8 = {llen | olen << 32}
        st64(&w.a8[s], ((uint64_t)tj << 32) | pslot);
        st64(&w.b8[s], ((uint64_t)__float_as_uint(olen) << 32) | __float_as_uint(llen));
        if constexpr (MODE == kSegment) {
#pragma unroll
          for (int t = 0; t < NT; ++t) std_(&w.acc[(size_t)s * NT + t], 0.0);
        }
      }
      // relax in-edges nb -> v; the atomic's old key also tells whether nb
      // settled before v (canonical edges count toward the circuit rank)
      const uint32_t r0 = a.rev_off[v], r1 = a.rev_off[v + 1];
      st_edges += (r1 - r0);
      for (uint32_t j = r0; j < r1; ++j) {
        const uint2 e = a.rev_adj[j];
        const uint32_t nb = e.x & kNodeMask;
        if (nb == v) continue;
        const float cs = secs_v + __uint_as_float(e.y);
        uint64_t kn = 0;
        if (cs <= a.maxsec) {
          const uint64_t nk = pack_key(c.tag, cs);
          kn = wmax64(&w.keyw[nb], nk);
          if (nk > kn) {
            const uint32_t b = bucket_of(cs, a.inv_delta);
            if (b <= c.k || b - c.k >= a.nbk) {
              lfail |= (b <= c.k) ? kFailInternal : kFailRing;
            } else {
              const uint32_t rb = b % a.nbk;
              const uint32_t pos = wadd32(&w.bcnt[rb], 1u);
              if (pos < a.bcap)
                st64(reinterpret_cast<uint64_t *>(&w.bkt[rb * a.bcap + pos]),
                     ((uint64_t)__float_as_uint(cs) << 32) | nb);
              else
                lfail |= kFailBucket;
              ++pushes;
            }
          }
        } else if (cycles && (e.x & kCanonBit)) {
          kn = ld64(&w.keyw[nb]);
        }
        if (cycles && (e.x & kCanonBit) && key_current(kn, c.tag) && key_less(key_secs(kn), nb, secs_v, v))
          ++cnt_e;
      }
      if (cycles) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (route_v <= a.dist[t]) {
            ++nband[t];
            eband[t] += cnt_e;
          }
        }
      }
    }
    pending -= cnt;
    pending += wave_sum(pushes);
    c.fail |= uniform(wave_or(lfail));
    wave_fence();
    if (c.fail) break;
    ++c.k;
  }
  if (c.fail) {
    // leave the ring empty for the next source
    for (uint32_t r = lane; r < a.nbk; r += kWave) st32(&w.bcnt[r], 0u);
    wave_fence();
    return false;
  }
  if (lane == 0) st32(&w.bnd[c.nb], c.nslot);
  wave_fence();
  return true;
}

// ---------------------------------------------------------------------------
// Persistent kernel: each wave loops over sources taken from a global counter.
// ---------------------------------------------------------------------------
template <int MODE, int NT>
__global__ __launch_bounds__(kBlock) void k_search(SearchArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wid = uniform((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
  const WaveWs w = wave_ws(a, wid);
  uint32_t epoch = a.epochs[wid];
  unsigned long long st_reached = 0, st_edges = 0, st_within = 0;
  const float speed = a.speed;
  for (;;) {
    uint32_t si = 0;
    if (lane == 0) si = atomicAdd(a.next_source, 1u);
    si = uniform(si);
    if (si >= a.n_sources) break;
    SrcCtx c;
    c.src = a.sources[si];
    ++epoch;
    c.tag = epoch;
    uint32_t nband[NT], eband[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) nband[t] = eband[t] = 0;
    if (!search_source<MODE, NT>(a, w, c, nband, eband, st_edges)) {
      if (lane == 0) {
        const uint32_t q = atomicAdd(a.n_fail, 1u);
        a.fail_list[q] = si;
        atomicOr(a.fail_flags, c.fail);
      }
      continue;
    }
    const uint32_t R = c.nslot;
    st_reached += (lane == 0) ? R : 0;

    if constexpr (MODE == kBrandes) {
      const float wt = a.src_wt[si];
      const float cyc = a.src_cyc[si];
      constexpr uint32_t M = 7;
      float rank[NT];
      uint32_t within[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t nn = wave_sum(nband[t]);
        const uint32_t ee = wave_sum(eband[t]);
        const int r = (int)ee - (int)nn + 1;
        rank[t] = (nn == 0) ? 0.0f : (float)(r > 0 ? r : 0);
        within[t] = 0;
      }
      // ---- closeness: scatter to targets (centrality.rs:1736-1779)
      for (uint32_t s = lane; s < R; s += kWave) {
        const uint32_t v = ld32(&w.node[s]);
        if (v == c.src) continue;
        const float route = ldf(&w.secs[s]) * speed;
        double *accv = a.acc + (size_t)v * M * NT;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (route <= a.dist[t]) {
            ++within[t];
            if (a.flags & 1u) {
              gadd_f64(&accv[0 * NT + t], (double)wt);
              gadd_f64(&accv[1 * NT + t], (double)(route * wt));
              gadd_f64(&accv[2 * NT + t], (double)(rank[t] * cyc));
              gadd_f64(&accv[3 * NT + t], (double)((1.0f / route) * wt));
              gadd_f64(&accv[4 * NT + t], (double)(expf(-a.beta[t] * route) * wt));
            }
          }
        }
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t tot = wave_sum(within[t]);
        st_within += (lane == 0) ? tot : 0;
        if ((a.flags & 5u) == 5u && lane == 0 && tot) atomicAdd(&a.reach[t], tot);
      }
      // ---- betweenness: reverse bucket order (centrality.rs:1783-1835, 823-873)
      if (a.flags & 2u) {
        for (int b = (int)c.nb - 1; b >= 0; --b) {
          const uint32_t s0 = ld32(&w.bnd[b]), s1 = ld32(&w.bnd[b + 1]);
          for (uint32_t s = s0 + lane; s < s1; s += kWave) {
            const uint32_t v = ld32(&w.node[s]);
            const float route = ldf(&w.secs[s]) * speed;
            const double sigma_w = ldd(reinterpret_cast<const double *>(&w.a8[s]));
            const uint64_t pm = ld64(&w.b8[s]);
            const bool is_src = v == c.src;
            const double pc = is_src ? 0.0 : (a.eligible[v] ? 0.5 : 1.0);
            const uint32_t f0 = a.fwd_off[v];
            double *accv = a.acc + (size_t)v * M * NT;
            double dep[NT], dep_b[NT], seed_b[NT];
            bool any = false;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              dep[t] = dep_b[t] = seed_b[t] = 0.0;
              if (route <= a.dist[t]) {
                seed_b[t] = pc * exp(-(double)a.beta[t] * (double)route);
                dep[t] = pc + ldd(&w.acc[(size_t)s * 2 * NT + t]);
                dep_b[t] = seed_b[t] + ldd(&w.acc[(size_t)s * 2 * NT + NT + t]);
                any = any || dep[t] != 0.0 || dep_b[t] != 0.0;
              }
            }
            if (!any) continue;
            for (uint64_t m = pm; m; m &= m - 1) {
              const uint32_t j = __builtin_ctzll(m);
              const uint32_t u = a.fwd_adj[f0 + j].x & kNodeMask;
              const uint32_t su = ld32(&w.slotw[u]);
              const double sigma_u = ldd(reinterpret_cast<const double *>(&w.a8[su]));
              if (sigma_u == 0.0) continue;
              const double fct = sigma_u / sigma_w;
#pragma unroll
              for (int t = 0; t < NT; ++t) {
                if (dep[t] != 0.0 || dep_b[t] != 0.0) {
                  wadd_f64(&w.acc[(size_t)su * 2 * NT + t], fct * dep[t]);
                  wadd_f64(&w.acc[(size_t)su * 2 * NT + NT + t], fct * dep_b[t]);
                }
              }
            }
            if (is_src) continue;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              if (dep[t] == 0.0 && dep_b[t] == 0.0) continue;
              const double credit = dep[t] - pc;
              const double credit_b = dep_b[t] - seed_b[t];
              if (credit > 0.0) gadd_f64(&accv[5 * NT + t], credit * (double)wt);
              if (credit_b > 0.0) gadd_f64(&accv[6 * NT + t], credit_b * (double)wt);
            }
          }
          wave_fence();
        }
      }
    } else if constexpr (MODE == kSegment) {
      constexpr uint32_t M = 4;
      // ---- closeness at the source over visited edges (centrality.rs:2201-2317)
      if (a.flags & 1u) {
        double sd[NT], sh[NT], sb[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) sd[t] = sh[t] = sb[t] = 0.0;
        for (uint32_t s = lane; s < R; s += kWave) {
          const uint32_t v = ld32(&w.node[s]);
          const float secs_v = ldf(&w.secs[s]);
          const float sd_n = secs_v * speed;
          const uint32_t r0 = a.rev_off[v], r1 = a.rev_off[v + 1];
          for (uint32_t j = r0; j < r1; ++j) {
            const uint32_t nb = a.rev_adj[j].x & kNodeMask;
            float sd_m = INFINITY;
            bool visited_edge = true;
            if (nb == v) {
              sd_m = sd_n;
            } else {
              const uint64_t kn = ld64(&w.keyw[nb]);
              if (key_current(kn, c.tag)) {
                const float s_nb = key_secs(kn);
                visited_edge = key_less(secs_v, v, s_nb, nb);
                sd_m = s_nb * speed;
              }
            }
            if (!visited_edge) continue;
            const float2 tw = a.rev_tw[j];
            const bool n_nearer = sd_n <= sd_m;
            const float av = n_nearer ? sd_n : sd_m;
            const float bv = n_nearer ? sd_m : sd_n;
            const float a_imp = av, b_imp = bv;
            const float elen = tw.x, imp = tw.y;
            const float cc = (elen + av + bv) / 2.0f;
            const float dd = cc;
            const float c_imp = a_imp + (cc - av) * imp;
            const float d_imp = c_imp;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              const float dist = a.dist[t];
              const float beta = a.beta[t];
              const float nbeta = -beta;
              const float inv = beta != 0.0f ? 1.0f / nbeta : 0.0f;
              if (av < dist) {
                float cur = cc, curi = c_imp;
                if (cur > dist) {
                  cur = dist;
                  curi = a_imp + (dist - av) * imp;
                }
                sd[t] += (double)(cur - av);
                const float hh = a_imp < 1.0f ? logf(curi) : logf(fmaxf(curi / a_imp, 1.1920929e-07f));
                sh[t] += (double)hh;
                const float bb = beta == 0.0f ? curi - a_imp : (expf(nbeta * curi) - expf(nbeta * a_imp)) * inv;
                sb[t] += (double)bb;
              }
              if (bv == dd) continue;
              if (bv <= dist) {
                float cur = dd, curi = d_imp;
                if (cur > dist) {
                  cur = dist;
                  curi = b_imp + (dist - bv) * imp;
                }
                sd[t] += (double)(cur - bv);
                const float hh = b_imp < 1.0f ? logf(curi) : logf(fmaxf(curi / b_imp, 1.1920929e-07f));
                sh[t] += (double)hh;
                const float bb = beta == 0.0f ? curi - b_imp : (expf(nbeta * curi) - expf(nbeta * b_imp)) * inv;
                sb[t] += (double)bb;
              }
            }
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const double t0 = wave_sum(sd[t]), t1 = wave_sum(sh[t]), t2 = wave_sum(sb[t]);
          if (lane == 0) {
            double *accs = a.acc + (size_t)c.src * M * NT;
            gadd_f64(&accs[0 * NT + t], t0);
            gadd_f64(&accs[1 * NT + t], t1);
            gadd_f64(&accs[2 * NT + t], t2);
          }
        }
      }
      // ---- betweenness: subtree sums along the single-pred tree (centrality.rs:2319-2401)
      if (a.flags & 2u) {
        for (int b = (int)c.nb - 1; b >= 0; --b) {
          const uint32_t s0 = ld32(&w.bnd[b]), s1 = ld32(&w.bnd[b + 1]);
          for (uint32_t s = s0 + lane; s < s1; s += kWave) {
            const uint32_t v = ld32(&w.node[s]);
            if (v == c.src) continue;
            const float sdist = ldf(&w.secs[s]) * speed;
            const uint32_t pslot = (uint32_t)ld64(&w.a8[s]);
            const uint64_t lb = ld64(&w.b8[s]);
            const float llen = __uint_as_float((uint32_t)lb);
            const float olen = __uint_as_float((uint32_t)(lb >> 32));
            const uint32_t pred = ld32(&w.node[pslot]);
            const float min_span = sdist - olen - llen;
            const float o_1 = min_span, o_2 = min_span + olen, l_1 = min_span, l_2 = min_span + llen;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              const double below = ldd(&w.acc[(size_t)s * NT + t]);
              if (below > 0.0) gadd_f64(&a.acc[(size_t)v * M * NT + 3 * NT + t], below);
              double own = 0.0;
              if (v > c.src) {
                const float dist = a.dist[t];
                const float beta = a.beta[t];
                if (min_span <= dist) {
                  const float o2s = fminf(o_2, dist), l2s = fminf(l_2, dist);
                  float auc;
                  if (beta == 0.0f) {
                    auc = (o2s - o_1) + (l2s - l_1);
                  } else {
                    const float nb2 = -beta, inv = 1.0f / nb2;
                    auc = (expf(nb2 * o2s) - expf(nb2 * o_1)) * inv + (expf(nb2 * l2s) - expf(nb2 * l_1)) * inv;
                  }
                  if (isfinite(auc) && auc >= 0.0f) own = (double)auc;
                }
              }
              const double up = below + own;
              if (pred != c.src && up != 0.0) wadd_f64(&w.acc[(size_t)pslot * NT + t], up);
            }
          }
          wave_fence();
        }
      }
    } else {
      // single-source exports (trees / Brandes state)
      for (uint32_t s = lane; s < R; s += kWave) {
        const uint32_t v = ld32(&w.node[s]);
        a.tout.secs[v] = ldf(&w.secs[s]);
        if constexpr (MODE == kTree) {
          const uint64_t pa = ld64(&w.a8[s]);
          if (v != c.src) {
            a.tout.pred[v] = (int32_t)ld32(&w.node[(uint32_t)pa]);
            a.tout.last_eid[v] = (int32_t)a.fwd_eid[a.fwd_off[v] + (uint32_t)(pa >> 32)];
          }
        } else if constexpr (MODE == kBrandesExport) {
          const uint64_t pm = ld64(&w.b8[s]);
          a.tout.sigma[v] = ldd(reinterpret_cast<const double *>(&w.a8[s]));
          uint32_t np = 0;
          const uint32_t f0 = a.fwd_off[v];
          for (uint64_t m = pm; m; m &= m - 1) {
            const uint32_t j = __builtin_ctzll(m);
            if (np < a.tout.max_preds)
              a.tout.pred_list[(size_t)v * a.tout.max_preds + np] = a.fwd_adj[f0 + j].x & kNodeMask;
            ++np;
          }
          a.tout.n_preds[v] = np;
        }
      }
      wave_fence();
    }
  }
  if (lane == 0) {
    a.epochs[wid] = epoch;
    atomicAdd(&a.stats[0], st_reached);
    atomicAdd(&a.stats[2], st_within);
  }
  st_edges = wave_sum(st_edges);
  if (lane == 0) atomicAdd(&a.stats[1], st_edges);
}

// ---------------------------------------------------------------------------
// Host side: workspace sizing and launch with retry.
// ---------------------------------------------------------------------------

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

template <int MODE>
void launch_nt(uint32_t T, uint32_t blocks, hipStream_t stream, const SearchArgs &a) {
  if constexpr (MODE == kTree || MODE == kBrandesExport) {
    hipLaunchKernelGGL((k_search<MODE, 1>), dim3(blocks), dim3(kBlock), 0, stream, a);
  } else {
    switch (T) {
#define CS_NT_CASE(n) \
  case n:             \
    hipLaunchKernelGGL((k_search<MODE, n>), dim3(blocks), dim3(kBlock), 0, stream, a); \
    break;
      CS_NT_CASE(1)
      CS_NT_CASE(2)
      CS_NT_CASE(3)
      CS_NT_CASE(4)
      CS_NT_CASE(5)
      CS_NT_CASE(6)
      CS_NT_CASE(7)
      CS_NT_CASE(8)
#undef CS_NT_CASE
      default:
        fail(CS_EINVAL, "unsupported number of thresholds");
    }
  }
}

struct Plan {
  float delta;
  uint32_t nbk, bcap, cap, maxb, waves;
  size_t stride;
  size_t o_keyw, o_slotw, o_node, o_secs, o_a8, o_b8, o_acc, o_bkt, o_bcnt, o_bnd;
};

void layout(Plan &p, uint32_t N, uint32_t T) {
  size_t o = 0;
  p.o_keyw = o;
  o = align256(o + 8ull * N);
  p.o_slotw = o;
  o = align256(o + 4ull * N);
  p.o_node = o;
  o = align256(o + 4ull * p.cap);
  p.o_secs = o;
  o = align256(o + 4ull * p.cap);
  p.o_a8 = o;
  o = align256(o + 8ull * p.cap);
  p.o_b8 = o;
  o = align256(o + 8ull * p.cap);
  p.o_acc = o;
  o = align256(o + 8ull * p.cap * 2 * T);
  p.o_bkt = o;
  o = align256(o + 8ull * p.nbk * p.bcap);
  p.o_bcnt = o;
  o = align256(o + 4ull * p.nbk);
  p.o_bnd = o;
  o = align256(o + 4ull * (p.maxb + 2));
  p.stride = o;
}

// Bucket width: below the smallest edge cost by the largest tie-band drift a
// node's fold can accumulate (one 1e-4 step per candidate edge), so candidates
// from the same bucket can never reach a receiver's tie band.
float choose_delta(const Graph &g, const DevEdges &de, float maxsec) {
  const float min_es = de.min_es;
  if (!(min_es > 0.0f)) fail(CS_EINVAL, "graph has zero-cost edges; the bucketed search needs positive edge costs");
  const double deg = (double)std::max(g.max_out_deg, g.max_in_deg);
  const double gap = (double)maxsec * 1.02e-4 * (deg + 3.0) + 1e-3;
  const double delta = (double)min_es - gap;
  if (delta < 0.05 * (double)min_es)
    fail(CS_EINVAL,
         "edge costs are too small relative to the distance threshold for the bucketed search "
         "(min edge seconds " + std::to_string(min_es) + ", max seconds " + std::to_string(maxsec) + ")");
  return (float)delta;
}

struct Workspace {
  int device = -1;
  char *ws = nullptr;
  size_t bytes = 0;
  uint32_t *epochs = nullptr;
  uint32_t waves = 0;
  uint64_t epoch_hi = 0;  // upper bound on epochs used
  uint32_t n = 0;
  size_t stride = 0;
  std::mutex mu;
};

std::mutex g_ws_mu;
std::map<int, std::vector<std::unique_ptr<Workspace>>> g_ws_pool;

Workspace *acquire_ws(int device) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto &v = g_ws_pool[device];
  for (auto &w : v)
    if (w->mu.try_lock()) return w.get();
  v.emplace_back(new Workspace());
  v.back()->device = device;
  v.back()->mu.lock();
  return v.back().get();
}

void release_ws(Workspace *w) { w->mu.unlock(); }

struct WsGuard {
  Workspace *w;
  ~WsGuard() { release_ws(w); }
};

// (Re)shape the workspace for `waves` waves of `stride` bytes over N nodes.
void prepare_ws(Workspace &w, uint32_t N, size_t stride, uint32_t waves, uint64_t sources,
                hipStream_t stream) {
  const size_t need = stride * waves;
  bool reinit = false;
  if (need > w.bytes || w.n != N || w.stride != stride || waves > w.waves) {
    if (w.ws) hip_check(hipFree(w.ws), "hipFree");
    if (w.epochs) hip_check(hipFree(w.epochs), "hipFree");
    w.ws = nullptr;
    w.epochs = nullptr;
    hipError_t e = hipMalloc(&w.ws, need);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      w.bytes = 0;
      fail(CS_ENOMEM, "could not allocate " + std::to_string(need >> 20) + " MiB of search workspace");
    }
    hip_check(hipMalloc(&w.epochs, 4ull * waves), "hipMalloc epochs");
    w.bytes = need;
    w.waves = waves;
    w.n = N;
    w.stride = stride;
    reinit = true;
  }
  if (w.epoch_hi + sources + 2 >= 0xfffffff0ull) reinit = true;
  if (reinit) {
    // zero memory: every key carries epoch 0 (never issued), every ring empty
    hip_check(hipMemsetAsync(w.ws, 0, need, stream), "hipMemset ws");
    hip_check(hipMemsetAsync(w.epochs, 0, 4ull * w.waves, stream), "hipMemset epochs");
    w.epoch_hi = 0;
  }
  w.epoch_hi += sources;
}

template <int MODE>
void run_search(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs, const uint32_t *d_sources,
                const float *d_wt, const float *d_cyc, uint32_t n_sources, const uint8_t *d_eligible,
                double *d_acc, uint32_t *d_reach, const TreeOut *tout, hipStream_t stream,
                cs_stats &st) {
  if (n_sources == 0) return;
  const uint32_t N = g.n;
  const uint32_t T = cs.T;
  Plan p;
  p.delta = choose_delta(g, de, cs.maxsec);
  const float inv_delta = 1.0f / p.delta;
  const uint32_t jump = (uint32_t)std::ceil((double)de.max_es / (double)p.delta) + 2;
  p.nbk = jump + 1;
  if (p.nbk > 4096) fail(CS_EINVAL, "edge cost spread too large for the bucket ring");
  p.maxb = (uint32_t)((double)cs.maxsec * (double)inv_delta) + 4;
  p.bcap = 1024;
  p.cap = std::min<uint32_t>(N, 16384);
  int dev = dg.device;
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  const uint32_t cus = (uint32_t)prop.multiProcessorCount;
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");

  // device scratch for the driver loop
  uint32_t *d_ctl = nullptr;  // next_source, n_fail, fail_flags, pad
  unsigned long long *d_stats = nullptr;
  uint32_t *d_fail = nullptr;
  uint32_t *d_src_cur = nullptr;
  float *d_wt_cur = nullptr, *d_cyc_cur = nullptr;
  hip_check(hipMalloc(&d_ctl, 16), "hipMalloc ctl");
  hip_check(hipMalloc(&d_stats, 32), "hipMalloc stats");
  hip_check(hipMalloc(&d_fail, 4ull * n_sources), "hipMalloc fail");
  hip_check(hipMemsetAsync(d_stats, 0, 32, stream), "memset stats");

  Workspace *wsp = acquire_ws(dev);
  WsGuard guard{wsp};
  const uint32_t *src_list = d_sources;
  const float *wt_list = d_wt;
  const float *cyc_list = d_cyc;
  uint32_t n_list = n_sources;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  double kms = 0.0;
  for (int attempt = 0; attempt < 6; ++attempt) {
    layout(p, N, T);
    // waves: enough to fill the chip, bounded by sources and by a quarter of free HBM
    uint32_t want = std::min<uint32_t>(n_list, cus * 32);
    const size_t budget = free_b / 4 + (wsp->bytes);
    uint32_t by_mem = (uint32_t)std::min<size_t>(budget / p.stride, 1u << 20);
    if (by_mem < 4) {
      if (p.cap > 4096) {
        p.cap = std::max<uint32_t>(4096, p.cap / 2);
        continue;
      }
      fail(CS_ENOMEM, "graph too large for the search workspace on this device");
    }
    uint32_t waves = std::min(want, by_mem);
    waves = std::max<uint32_t>(4, (waves + 3) / 4 * 4);
    p.waves = waves;
    prepare_ws(*wsp, N, p.stride, waves, n_list, stream);
    hip_check(hipMemsetAsync(d_ctl, 0, 16, stream), "memset ctl");
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = N;
    a.rev_off = dg.rev_off;
    a.rev_adj = de.rev_adj;
    a.fwd_off = dg.fwd_off;
    a.fwd_adj = de.fwd_adj;
    a.rev_tw = de.rev_tw;
    a.fwd_twlen = de.fwd_twlen;
    a.fwd_eid = de.fwd_eid;
    a.eligible = d_eligible;
    a.T = T;
    a.maxsec = cs.maxsec;
    a.speed = cs.speed;
    a.inv_delta = inv_delta;
    a.tolerance = cs.tolerance;
    for (uint32_t i = 0; i < kMaxT; ++i) {
      a.dist[i] = i < T ? cs.dist[i] : 0.0f;
      a.beta[i] = i < T ? cs.beta[i] : 0.0f;
    }
    a.nbk = p.nbk;
    a.bcap = p.bcap;
    a.cap = p.cap;
    a.maxb = p.maxb;
    a.n_sources = n_list;
    a.sources = src_list;
    a.src_wt = wt_list;
    a.src_cyc = cyc_list;
    a.flags = (cs.closeness ? 1u : 0u) | (cs.betweenness ? 2u : 0u) | (cs.track_reach ? 4u : 0u);
    a.acc = d_acc;
    a.reach = d_reach;
    a.stats = d_stats;
    a.ws = wsp->ws;
    a.ws_stride = p.stride;
    a.o_keyw = p.o_keyw;
    a.o_slotw = p.o_slotw;
    a.o_node = p.o_node;
    a.o_secs = p.o_secs;
    a.o_a8 = p.o_a8;
    a.o_b8 = p.o_b8;
    a.o_acc = p.o_acc;
    a.o_bkt = p.o_bkt;
    a.o_bcnt = p.o_bcnt;
    a.o_bnd = p.o_bnd;
    a.epochs = wsp->epochs;
    a.next_source = d_ctl;
    a.n_fail = d_ctl + 1;
    a.fail_flags = d_ctl + 2;
    a.fail_list = d_fail;
    if (tout) a.tout = *tout;
    {
      const uint32_t blocks = waves / 4;
      hip_check(hipEventRecord(e0, stream), "event record");
      launch_nt<MODE>(T, blocks, stream, a);
      hip_check(hipGetLastError(), "k_search launch");
      hip_check(hipEventRecord(e1, stream), "event record");
    }
    uint32_t ctl[4];
    hip_check(hipMemcpyAsync(ctl, d_ctl, 16, hipMemcpyDeviceToHost, stream), "copy ctl");
    hip_check(hipStreamSynchronize(stream), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    kms += ms;
    st.waves = waves;
    const uint32_t n_failed = ctl[1], flags = ctl[2];
    if (n_failed == 0) break;
    if (flags & kFailInternal) fail(CS_EINTERNAL, "search invariant violated (bucket order)");
    if (flags & kFailRing) fail(CS_EINTERNAL, "bucket ring window exceeded");
    st.retries += n_failed;
    // retry the failed sources with a larger workspace
    if (!d_src_cur) {
      hip_check(hipMalloc(&d_src_cur, 4ull * n_sources), "hipMalloc");
      hip_check(hipMalloc(&d_wt_cur, 4ull * n_sources), "hipMalloc");
      hip_check(hipMalloc(&d_cyc_cur, 4ull * n_sources), "hipMalloc");
    }
    // gather failed (source, wt, cyc) on the host: small lists
    std::vector<uint32_t> fl(n_failed), srcs(n_list);
    std::vector<float> wts(n_list, 0.0f), cycs(n_list, 0.0f);
    hip_check(hipMemcpy(fl.data(), d_fail, 4ull * n_failed, hipMemcpyDeviceToHost), "copy");
    hip_check(hipMemcpy(srcs.data(), src_list, 4ull * n_list, hipMemcpyDeviceToHost), "copy");
    if (wt_list) hip_check(hipMemcpy(wts.data(), wt_list, 4ull * n_list, hipMemcpyDeviceToHost), "copy");
    if (cyc_list) hip_check(hipMemcpy(cycs.data(), cyc_list, 4ull * n_list, hipMemcpyDeviceToHost), "copy");
    std::vector<uint32_t> ns(n_failed);
    std::vector<float> nw(n_failed), nc(n_failed);
    for (uint32_t i = 0; i < n_failed; ++i) {
      ns[i] = srcs[fl[i]];
      nw[i] = wts[fl[i]];
      nc[i] = cycs[fl[i]];
    }
    hip_check(hipMemcpy(d_src_cur, ns.data(), 4ull * n_failed, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(d_wt_cur, nw.data(), 4ull * n_failed, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(d_cyc_cur, nc.data(), 4ull * n_failed, hipMemcpyHostToDevice), "copy");
    src_list = d_src_cur;
    wt_list = wt_list ? d_wt_cur : nullptr;
    cyc_list = cyc_list ? d_cyc_cur : nullptr;
    n_list = n_failed;
    if (flags & kFailCap) p.cap = std::min<uint32_t>(N, p.cap * 4);
    if (flags & kFailBucket) p.bcap *= 4;
    if (attempt == 5) fail(CS_EINTERNAL, "search workspace retries exhausted");
  }
  unsigned long long stv[4];
  hip_check(hipMemcpy(stv, d_stats, 32, hipMemcpyDeviceToHost), "copy stats");
  st.kernel_ms += kms;
  st.nodes_reached += stv[0];
  st.edges_scanned += stv[1];
  st.within_sum += stv[2];
  hip_check(hipEventDestroy(e0), "event");
  hip_check(hipEventDestroy(e1), "event");
  hip_check(hipFree(d_ctl), "free");
  hip_check(hipFree(d_stats), "free");
  hip_check(hipFree(d_fail), "free");
  if (d_src_cur) {
    hip_check(hipFree(d_src_cur), "free");
    hip_check(hipFree(d_wt_cur), "free");
    hip_check(hipFree(d_cyc_cur), "free");
  }
}

}  // namespace

void launch_shortest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                     const uint32_t *d_sources, const float *d_wt, const float *d_cyc,
                     uint32_t n_sources, const uint8_t *d_eligible, double *d_acc, uint32_t *d_reach,
                     hipStream_t stream, cs_stats &st) {
  run_search<kBrandes>(g, dg, de, cs, d_sources, d_wt, d_cyc, n_sources, d_eligible, d_acc, d_reach,
                       nullptr, stream, st);
}

void launch_segment(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                    const uint32_t *d_sources, uint32_t n_sources, double *d_acc, hipStream_t stream,
                    cs_stats &st) {
  run_search<kSegment>(g, dg, de, cs, d_sources, nullptr, nullptr, n_sources, nullptr, d_acc, nullptr,
                       nullptr, stream, st);
}

void launch_single(Graph &g, DevGraph &dg, DevEdges &de, uint32_t src, float maxsec, float tolerance,
                   int mode, const TreeOut &out, hipStream_t stream) {
  CallSpec cs;
  cs.T = 1;
  cs.dist[0] = maxsec * de.speed;
  cs.maxsec = maxsec;
  cs.speed = de.speed;
  cs.tolerance = tolerance;
  cs.closeness = false;
  cs.betweenness = false;
  uint32_t *d_src = nullptr;
  hip_check(hipMalloc(&d_src, 4), "hipMalloc");
  hip_check(hipMemcpy(d_src, &src, 4, hipMemcpyHostToDevice), "copy");
  cs_stats st{};
  if (mode == 0)
    run_search<kTree>(g, dg, de, cs, d_src, nullptr, nullptr, 1, nullptr, nullptr, nullptr, &out, stream, st);
  else
    run_search<kBrandesExport>(g, dg, de, cs, d_src, nullptr, nullptr, 1, nullptr, nullptr, nullptr, &out,
                               stream, st);
  hip_check(hipFree(d_src), "free");
}

}  // namespace cs
