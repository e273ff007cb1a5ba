// cs_angular.hip -- simplest-path (angular) centrality on the dual graph.
//
// Angular route increments are turn angles plus 1e-6 x length
// (centrality.rs:665-667): straight continuations cost almost nothing, so no
// bucket width separates a state from its predecessors and the reference's
// order-dependent fold (centrality.rs:613-705) must see candidates strictly in
// settle order. Each lane therefore runs one source's search serially with a
// private priority queue keyed (route, seconds at push, state): a ring of
// buckets on route cost, or a 4-ary heap for a source that overfills a bucket.
// Up to 64 independent sources run per wave, one per lane, dealt statically.
//
// Per-lane state: epoch-tagged maps state -> slot and node -> slot (zero
// memory is stale, nothing cleared per source), slot arrays for reached
// states and nodes, the heap, and the settle order. Slot arrays and the heap
// are interleaved across the 64 lanes of a wave so that lanes stepping in
// lockstep (heap top, slot allocation) touch the same cache lines.
#include "cs_internal.hpp"

#include <cmath>
#include <cstring>
#include <type_traits>

namespace cs {

namespace {

constexpr uint32_t kFailCap = 1u, kFailHeap = 2u, kFailPreds = 4u;
constexpr uint32_t kNoSource = 0xffffffffu;  // padding item
constexpr uint32_t kRelax = 4;               // out-edges whose loads one relaxation step batches
constexpr uint32_t kNoEdge = 0xffffffffu;     // padding entry of a state row
constexpr uint32_t kRowSpill = 0xfffffffeu;   // first entry of a state row whose edges do not fit

// One (graph, request) pair of a launch: a plain call has one, a served slice
// one per active request.
struct AngJob {
  uint32_t N;           // nodes of the graph (accumulator rows)
  const uint32_t *off;  // out-edge rows (petgraph order)
  const uint4 *adj;     // {v | src_slot<<30 | dst_slot<<31, bits(es), bits(angle), bits(1e-6*len)}
  const uint4 *srow;    // [2N][kRelax] each state's usable out-edges in row order (kNoEdge pads; kRowSpill: use the row)
  const uint8_t *eligible;
  uint32_t T;
  float maxsec, tolerance, unit, offset;
  float sec_thr[kMaxT];
  const uint32_t *sources;
  const float *src_wt;
  uint32_t flags;  // 1 closeness, 2 betweenness, 4 track reach
  double *acc;     // [4][tstride][N], metric-major
  float bq_inv;    // bucket queue: buckets per unit of route (the ring spans more than the largest edge)
  uint32_t tstride;  // threshold stride of acc rows (>= T)
  uint32_t *reach;
};

// Jobs are read-only for the whole launch: reading them through the constant
// address space lets the compiler keep their fields in registers across the
// workspace stores instead of reloading them.
typedef __attribute__((address_space(4))) const AngJob CAngJob;

struct AngArgs {
  uint32_t N;           // largest graph of the launch (lane map layout)
  const AngJob *jobs;
  const uint2 *items;   // {job, source position}; NULL: job 0, position = item
  uint32_t n_items;
  unsigned long long *stats;
  char *ws;
  size_t lane_stride;  // bytes of the lane-major maps per lane
  size_t wave_stride;  // bytes of the interleaved arrays per wave
  uint32_t scap, ncap, hcap, maxp, tcap;  // per-lane capacities (tcap: thresholds of the delta/seed blocks)
  size_t o_srec, o_preds, o_order, o_orte, o_delta, o_seed, o_nrec, o_heap;
  size_t o_bqe;  // bucket queue entries [kBuckets][bq_cap] per lane (the counts live in LDS)
  uint32_t bq_cap;
  char *wave_ws;
  uint32_t per_wave;    // lanes of a wave that take sources (the rest take padding items)
  uint32_t epoch_base;  // item it searches under epoch epoch_base + it + 1
  uint32_t *epochs;     // single-source tree: its one epoch counter
  uint32_t *next_source;
  uint32_t *fail_list;
  uint32_t *n_fail;
  uint32_t *fail_flags;
  // single-source tree (dijkstra_tree_simplest) outputs, N long
  float *t_simpl;
  float *t_secs;
  int32_t *t_pred;
  uint8_t *t_vis;
  uint32_t *t_order;
  uint32_t *t_norder;
  uint32_t *progress;  // [lanes] sources each lane has finished (pinned host memory), or null
};

// Element i of this lane's own block in a wave-shared array of 64 lane blocks
// of `count` elements. Lanes search different sources at unrelated indices, so
// nothing coalesces across lanes; a lane's neighbouring elements (the four
// children of a heap node, consecutive settle positions) share a cache line.
template <typename T>
__device__ __forceinline__ T *il(char *base, size_t off, uint32_t i, uint32_t lane, size_t count) {
  return reinterpret_cast<T *>(base + off) + (size_t)lane * count + i;
}

// A search state's fields live together in one 32-byte record, a node's in
// one 16-byte record, so that the fields a step reads or writes for a slot
// share one cache line (records never straddle one).
constexpr uint32_t kSRec = 32, kNRec = 16;
constexpr uint32_t kS_route = 0;   // f32 route metric
constexpr uint32_t kS_secs = 4;    // f32 seconds
constexpr uint32_t kS_state = 8;   // u32 state id, visited flag in bit 31
constexpr uint32_t kS_np = 12;     // u32 predecessor count
constexpr uint32_t kS_sigma = 16;  // f64 path count
constexpr uint32_t kN_node = 0;    // u32 node id
constexpr uint32_t kN_brt = 4;     // f32 best route over the node's states
constexpr uint32_t kN_bsec = 8;    // f32 seconds of that best route
template <typename T>
__device__ __forceinline__ T *sr(char *base, size_t off, uint32_t i, uint32_t lane, size_t count, uint32_t f) {
  return reinterpret_cast<T *>(base + off + ((size_t)lane * count + i) * kSRec + f);
}
template <typename T>
__device__ __forceinline__ T *nr(char *base, size_t off, uint32_t i, uint32_t lane, size_t count, uint32_t f) {
  return reinterpret_cast<T *>(base + off + ((size_t)lane * count + i) * kNRec + f);
}

struct LaneWs {
  uint64_t *skey;  // [2N] epoch<<32 | state slot
  uint64_t *nkey;  // [N]  epoch<<32 | node slot
  char *w;         // wave-interleaved arrays
  uint32_t lane;
};

__device__ __forceinline__ bool hless(float ka, float sa, uint32_t ia, float kb, float sb, uint32_t ib) {
  if (ka != kb) return ka < kb;
  if (sa != sb) return sa < sb;
  return ia < ib;
}

// Per-lane 4-ary min-heap on (route, seconds at push, state). Every key is
// distinct (a state is pushed again only with a strictly smaller route), so the
// pop sequence is the key order whatever the heap's shape; four children per
// level halve the depth of the dependent loads a pop walks down.
struct Heap {
  char *w;
  size_t off;
  uint32_t lane;
  uint32_t n;
  uint32_t cap;
  __device__ __forceinline__ uint4 *at(uint32_t i) { return il<uint4>(w, off, i, lane, cap); }
  __device__ __forceinline__ static bool less(const uint4 &a, const uint4 &b) {
    return hless(__uint_as_float(a.x), __uint_as_float(a.y), a.z, __uint_as_float(b.x), __uint_as_float(b.y), b.z);
  }
  // `slot` rides along outside the key (the state's workspace slot); false: full
  __device__ __forceinline__ bool push(float k, float s, uint32_t idx, uint32_t slot = 0u) {
    if (n >= cap) return false;
    const uint4 e = make_uint4(__float_as_uint(k), __float_as_uint(s), idx, slot);
    uint32_t i = n++;
    while (i > 0) {
      const uint32_t p = (i - 1) / 4;
      const uint4 pe = *at(p);
      if (!less(e, pe)) break;
      *at(i) = pe;
      i = p;
    }
    *at(i) = e;
    return true;
  }
  __device__ __forceinline__ uint4 pop() {
    const uint4 top = *at(0);
    const uint4 last = *at(--n);
    uint32_t i = 0;
    for (;;) {
      const uint32_t c0 = 4 * i + 1;
      if (c0 >= n) break;
      // the (up to) four children are independent loads
      uint4 ch[4];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) ch[q] = (c0 + q < n) ? *at(c0 + q) : make_uint4(0u, 0u, 0u, 0u);
      uint32_t sidx = c0;
      uint4 best = ch[0];
#pragma unroll
      for (uint32_t q = 1; q < 4; ++q)
        if (c0 + q < n && less(ch[q], best)) {
          sidx = c0 + q;
          best = ch[q];
        }
      if (!less(best, last)) break;
      *at(i) = best;
      i = sidx;
    }
    if (n) *at(i) = last;
    return top;
  }
};

// Per-lane monotone bucket queue on route cost, the same interface and pop
// order as Heap. A search's keys never fall below the last key popped (edge
// costs are non-negative and f32 addition is monotone), so entries live in a
// ring of kBuckets buckets of width 1 / inv, a bucket holding every key whose
// floor(route * inv) maps to it; the host makes kBucketSpan buckets span the
// largest edge cost, so a pushed key lands fewer than kBuckets buckets past the
// last pop. A mask in registers marks the non-empty buckets. A pop scans the lowest
// non-empty bucket for its exact minimum under the heap's key order (route,
// seconds at push, state) and moves the bucket's last entry into the gap: keys
// are distinct (a state is pushed again only with a smaller route), so the pop
// sequence equals the heap's. Pops cost a count load and one batch of entry
// loads instead of a dependent load per heap level. A push that would leave
// the ring or overfill its bucket returns false (the source is searched again
// with the heap).
constexpr uint32_t kMaskWords = 2;                // occupancy mask words held in registers
constexpr uint32_t kBuckets = 64 * kMaskWords;     // ring of 128 buckets
constexpr uint32_t kBucketSpan = kBuckets - 8;     // buckets per largest route increment
constexpr uint32_t kBqCap = 64;  // entries per bucket before a source falls back to the heap
constexpr uint32_t kCloseBatch = 4;  // node records the closeness pass loads ahead of its atomics
// Entries per bucket: kBqCap, or 1/128 of a tuned slot capacity (at least 2),
// so that a small capacity also sends sources to the heap early.
uint32_t bucket_capacity() {
  const uint32_t cap0 = default_tuning().cap;
  return cap0 ? std::min<uint32_t>(kBqCap, std::max<uint32_t>(cap0 / 128, 2u)) : kBqCap;
}

// Bucket entry counts of the workgroup's lanes, in LDS: [bucket][lane], one
// byte each (a bucket holds at most 64 entries). A pop and a push read the
// count before they touch the bucket, so it sits on each lane's dependent
// chain; in LDS that read costs no memory round trip.
static_assert(kBqCap <= 255, "bucket counts are one byte");
__device__ __forceinline__ uint8_t *bq_counts() {
  __shared__ uint8_t c[kBuckets * kBlock];
  return c + threadIdx.x;
}

struct BQueue {
  char *w;
  size_t off_e;
  uint8_t *lc;  // this lane's column of the LDS counts
  uint32_t lane, cap;  // entries per bucket
  float inv;
  uint64_t mask[kMaskWords] = {};
  uint32_t cur = 0;  // absolute bucket of the last pop
  uint32_t n = 0;
  __device__ __forceinline__ uint4 *ent(uint32_t r, uint32_t i) {
    return il<uint4>(w, off_e, r * cap + i, lane, (size_t)kBuckets * cap);
  }
  __device__ __forceinline__ uint8_t *cnt(uint32_t r) { return lc + r * kBlock; }
  // mask words are indexed through selects so that they stay in registers
  __device__ __forceinline__ uint64_t word(uint32_t i) const {
    uint64_t v = 0;
#pragma unroll
    for (uint32_t k = 0; k < kMaskWords; ++k) v = (i == k) ? mask[k] : v;
    return v;
  }
  __device__ __forceinline__ void flip(uint32_t r, bool on) {
    const uint32_t i = r >> 6;
    const uint64_t b = 1ull << (r & 63);
#pragma unroll
    for (uint32_t k = 0; k < kMaskWords; ++k)
      if (i == k) mask[k] = on ? (mask[k] | b) : (mask[k] & ~b);
  }
  // distance from the last pop's bucket to the next non-empty one (n > 0)
  __device__ __forceinline__ uint32_t next_offset() const {
    const uint32_t c0 = cur & (kBuckets - 1), w0 = c0 >> 6, b0 = c0 & 63;
    const uint64_t x = word(w0) >> b0;
    if (x) return (uint32_t)__builtin_ctzll(x);
    uint32_t off = 64 - b0;
#pragma unroll
    for (uint32_t k = 1; k <= kMaskWords; ++k) {
      const uint64_t y = word((w0 + k) & (kMaskWords - 1));
      if (y) return off + (uint32_t)__builtin_ctzll(y);
      off += 64;
    }
    return 0;
  }
  __device__ __forceinline__ bool push(float k, float s, uint32_t idx, uint32_t slot = 0u) {
    const float fb = k * inv;
    if (!(fb < 4.0e9f)) return false;
    const uint32_t b = (uint32_t)fb;  // floor: k >= 0
    if (b - cur >= kBuckets) return false;
    const uint32_t r = b & (kBuckets - 1);
    const bool used = (word(r >> 6) >> (r & 63)) & 1ull;
    const uint32_t m = used ? *cnt(r) : 0u;
    if (m >= cap) return false;
    *ent(r, m) = make_uint4(__float_as_uint(k), __float_as_uint(s), idx, slot);
    *cnt(r) = (uint8_t)(m + 1);
    flip(r, true);
    ++n;
    return true;
  }
  __device__ __forceinline__ uint4 pop() {
    cur += next_offset();
    const uint32_t r = cur & (kBuckets - 1);
    const uint32_t m = *cnt(r);
    uint4 best = make_uint4(0u, 0u, 0u, 0u), last = best;
    uint32_t bi = 0;
    // entries in batches of eight independent loads
    for (uint32_t i0 = 0; i0 < m; i0 += 8) {
      uint4 e[8];
#pragma unroll
      for (uint32_t q = 0; q < 8; ++q) e[q] = (i0 + q < m) ? *ent(r, i0 + q) : make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
      for (uint32_t q = 0; q < 8; ++q) {
        const uint32_t i = i0 + q;
        if (i >= m) continue;
        if (i == 0 || Heap::less(e[q], best)) {
          best = e[q];
          bi = i;
        }
        if (i == m - 1) last = e[q];
      }
    }
    if (bi != m - 1) *ent(r, bi) = last;
    *cnt(r) = (uint8_t)(m - 1);
    if (m == 1) flip(r, false);
    --n;
    return best;
  }
};

// The lane's queue for a search: the bucket ring (QK 1) or the heap.
template <int QK>
__device__ __forceinline__ auto make_queue(const AngArgs &a, CAngJob &jb, char *w, uint32_t lane) {
  if constexpr (QK == 1) {
    return BQueue{w, a.o_bqe, bq_counts(), lane, a.bq_cap, jb.bq_inv};
  } else {
    return Heap{w, a.o_heap, lane, 0, a.hcap};
  }
}

// One source, exactly as dijkstra_brandes_angular Phase 1 (centrality.rs:577-705).
// Returns false (with fail bits) if a workspace capacity is exceeded.
template <typename Q>
__device__ bool angular_source(const AngArgs &a, CAngJob &jb, const LaneWs &L, Q &h, uint32_t src, uint32_t epoch,
                               uint32_t &n_states, uint32_t &n_nodes, uint32_t &n_order, uint32_t &fail) {
  const float one_m = 1.0f - kTieEps, one_p = 1.0f + kTieEps;
  char *w = L.w;
  const uint32_t lane = L.lane;
  n_states = n_nodes = n_order = 0;
  // node slot of the source
  L.nkey[src] = ((uint64_t)epoch << 32) | 0u;
  *nr<uint32_t>(w, a.o_nrec, 0, lane, a.ncap, kN_node) = src;
  *nr<float>(w, a.o_nrec, 0, lane, a.ncap, kN_brt) = 0.0f;
  *nr<float>(w, a.o_nrec, 0, lane, a.ncap, kN_bsec) = 0.0f;
  n_nodes = 1;
  for (uint32_t slot = 0; slot < 2; ++slot) {
    const uint32_t s = 2 * src + slot;
    const uint32_t ss = n_states++;
    L.skey[s] = ((uint64_t)epoch << 32) | ss;
    *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_route) = 0.0f;
    *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_secs) = 0.0f;
    *sr<double>(w, a.o_srec, ss, lane, a.scap, kS_sigma) = 1.0;
    *sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_state) = s;  // visited flag in bit 31
    *sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_np) = 0;
    h.push(0.0f, 0.0f, s, ss);
  }
  while (h.n > 0) {
    const uint4 top = h.pop();
    const uint32_t us = top.z;
    const uint32_t uss = top.w;  // the state's slot, carried by the heap entry
    const uint32_t u = us >> 1, entry = us & 1u;
    // the state's fields and its out-row bounds are independent loads
    uint32_t *ust = sr<uint32_t>(w, a.o_srec, uss, lane, a.scap, kS_state);
    const uint4 uhd = *sr<uint4>(w, a.o_srec, uss, lane, a.scap, 0);  // route, secs, state, np
    const uint32_t ustate = uhd.z;
    const float u_route = __uint_as_float(uhd.x);
    const float u_secs = __uint_as_float(uhd.y);
    const double u_sigma = *sr<double>(w, a.o_srec, uss, lane, a.scap, kS_sigma);
    // the state's usable out-edges: its state row (independent of the record
    // load), or for a state with more than kRelax of them the node's out-row
    uint4 trow[kRelax];
#pragma unroll
    for (uint32_t q = 0; q < kRelax; ++q) trow[q] = jb.srow[(size_t)us * kRelax + q];
    const bool from_row = trow[0].x == kRowSpill;
    const uint32_t j0 = from_row ? jb.off[u] : 0u, j1 = from_row ? jb.off[u + 1] : kRelax;
    if (ustate & 0x80000000u) continue;
    *ust = us | 0x80000000u;
    if (n_order >= a.scap) {
      fail |= kFailCap;
      return false;
    }
    *il<float>(w, a.o_orte, n_order, lane, a.scap) = u_route;  // final: a settled state is never relaxed
    *il<uint32_t>(w, a.o_order, n_order++, lane, a.scap) = uss;
    // out-edges in chunks of kRelax: the adjacency entries, then the state and
    // node keys, then the state and node fields are loaded for the whole chunk
    // at once; the edges are then applied one by one in row order, re-reading
    // a state or node that an earlier edge of the chunk has just written
    for (uint32_t jc = j0; jc < j1; jc += kRelax) {
      uint4 e[kRelax];
#pragma unroll
      for (uint32_t q = 0; q < kRelax; ++q)
        e[q] = !from_row ? trow[q] : (jc + q < j1) ? jb.adj[jc + q] : make_uint4(0u, 0u, 0u, 0u);
      bool ok[kRelax];
      uint32_t vv[kRelax], vsv[kRelax];
      float csv[kRelax];
      uint64_t skv[kRelax], nkv[kRelax];
#pragma unroll
      for (uint32_t q = 0; q < kRelax; ++q) {
        vv[q] = e[q].x & 0x3fffffffu;
        vsv[q] = 2 * vv[q] + (e[q].x >> 31);
        csv[q] = u_secs + __uint_as_float(e[q].y);
        ok[q] = (from_row ? (jc + q < j1 && ((e[q].x >> 30) & 1u) == 1u - entry) : e[q].x != kNoEdge) &&
                !(csv[q] > jb.maxsec);
        skv[q] = ok[q] ? L.skey[vsv[q]] : 0ull;
        nkv[q] = ok[q] ? L.nkey[vv[q]] : 0ull;
      }
      uint32_t flv[kRelax];
      float rtv[kRelax], brv[kRelax], bsv[kRelax];
#pragma unroll
      for (uint32_t q = 0; q < kRelax; ++q) {
        flv[q] = 0u;
        rtv[q] = brv[q] = bsv[q] = INFINITY;
        if (ok[q] && (uint32_t)(skv[q] >> 32) == epoch) {
          const uint4 r = *sr<uint4>(w, a.o_srec, (uint32_t)skv[q], lane, a.scap, 0);
          flv[q] = r.z;
          rtv[q] = __uint_as_float(r.x);
        }
        if (ok[q] && (uint32_t)(nkv[q] >> 32) == epoch) {
          const uint4 r = *nr<uint4>(w, a.o_nrec, (uint32_t)nkv[q], lane, a.ncap, 0);
          brv[q] = __uint_as_float(r.y);
          bsv[q] = __uint_as_float(r.z);
        }
      }
#pragma unroll
      for (uint32_t q = 0; q < kRelax; ++q) {
        if (!ok[q]) continue;
        bool dup_s = false, dup_n = false;
#pragma unroll
        for (uint32_t r = 0; r < q; ++r) {
          dup_s = dup_s || (ok[r] && vsv[r] == vsv[q]);
          dup_n = dup_n || (ok[r] && vv[r] == vv[q]);
        }
        const uint32_t v = vv[q], vs = vsv[q];
        const float cs = csv[q];
        uint64_t sk = skv[q];
        uint32_t fl = flv[q];
        float rt = rtv[q];
        if (dup_s) {
          sk = L.skey[vs];
          if ((uint32_t)(sk >> 32) == epoch) {
            fl = *sr<uint32_t>(w, a.o_srec, (uint32_t)sk, lane, a.scap, kS_state);
            rt = *sr<float>(w, a.o_srec, (uint32_t)sk, lane, a.scap, kS_route);
          }
        }
        uint32_t vss;
        float cur;
        if ((uint32_t)(sk >> 32) == epoch) {
          vss = (uint32_t)sk;
          if (fl & 0x80000000u) continue;  // visited
          cur = rt;
        } else {
          if (n_states >= a.scap) {
            fail |= kFailCap;
            return false;
          }
          vss = n_states++;
          L.skey[vs] = ((uint64_t)epoch << 32) | vss;
          *sr<uint4>(w, a.o_srec, vss, lane, a.scap, 0) =
              make_uint4(__float_as_uint(INFINITY), __float_as_uint(INFINITY), vs, 0u);
          *sr<double>(w, a.o_srec, vss, lane, a.scap, kS_sigma) = 0.0;
          cur = INFINITY;
        }
        float cr = u_route + __uint_as_float(e[q].z);
        cr = cr + __uint_as_float(e[q].w);
        uint32_t *np = sr<uint32_t>(w, a.o_srec, vss, lane, a.scap, kS_np);
        uint32_t *preds = reinterpret_cast<uint32_t *>(w + a.o_preds) + ((size_t)lane * a.scap + vss) * a.maxp;
        if (cr < cur) {
          double *sg = sr<double>(w, a.o_srec, vss, lane, a.scap, kS_sigma);
          // a clear improvement restarts the predecessor list without reading it back
          uint32_t k;
          if (cr < cur * one_m) {
            k = 0;
            *sg = u_sigma;
          } else {
            k = *np;
            *sg = *sg + u_sigma;
          }
          *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_route) = cr;
          *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_secs) = cs;
          if (k >= a.maxp) {
            fail |= kFailPreds;
            return false;
          }
          preds[k] = uss;
          *np = k + 1;
          if (!h.push(cr, cs, vs, vss)) {
            fail |= kFailHeap;
            return false;
          }
        } else if (cr <= cur * one_p) {
          const uint32_t k = *np;
          bool has = false;
          for (uint32_t r = 0; r < k; ++r)
            if (preds[r] == uss) has = true;
          if (!has) {
            float *vsec = sr<float>(w, a.o_srec, vss, lane, a.scap, kS_secs);
            if (cs < *vsec) *vsec = cs;
            if (k >= a.maxp) {
              fail |= kFailPreds;
              return false;
            }
            preds[k] = uss;
            *np = k + 1;
            double *sg = sr<double>(w, a.o_srec, vss, lane, a.scap, kS_sigma);
            *sg = *sg + u_sigma;
          }
        }
        // per-node best route / seconds (centrality.rs:695-703)
        uint64_t nk = nkv[q];
        float best = brv[q], bsec = bsv[q];
        if (dup_n) {
          nk = L.nkey[v];
          if ((uint32_t)(nk >> 32) == epoch) {
            best = *nr<float>(w, a.o_nrec, (uint32_t)nk, lane, a.ncap, kN_brt);
            bsec = *nr<float>(w, a.o_nrec, (uint32_t)nk, lane, a.ncap, kN_bsec);
          }
        }
        uint32_t nsl;
        if ((uint32_t)(nk >> 32) == epoch) {
          nsl = (uint32_t)nk;
        } else {
          if (n_nodes >= a.ncap) {
            fail |= kFailCap;
            return false;
          }
          nsl = n_nodes++;
          L.nkey[v] = ((uint64_t)epoch << 32) | nsl;
          *nr<uint32_t>(w, a.o_nrec, nsl, lane, a.ncap, kN_node) = v;
          best = bsec = INFINITY;
        }
        if (cr < best * one_m) {
          *nr<float>(w, a.o_nrec, nsl, lane, a.ncap, kN_brt) = cr;
          *nr<float>(w, a.o_nrec, nsl, lane, a.ncap, kN_bsec) = cs;
        } else if (cr <= best * one_p) {
          *nr<float>(w, a.o_nrec, nsl, lane, a.ncap, kN_bsec) = fminf(bsec, cs);
        } else if (!((uint32_t)(nk >> 32) == epoch)) {
          // a new node slot keeps its initial infinite best
          *nr<float>(w, a.o_nrec, nsl, lane, a.ncap, kN_brt) = INFINITY;
          *nr<float>(w, a.o_nrec, nsl, lane, a.ncap, kN_bsec) = INFINITY;
        }
      }
    }
  }
  return true;
}

// Phase 2 of the angular fold (centrality.rs:705-764): with a tolerance above
// the internal tie band, predecessors and sigma are rebuilt from the final
// routes, visiting settled states in settle order: a successor state settled
// later gains u as a predecessor when route_u + angle + 1e-6 length <=
// route_v * (1 + tol). Settle positions live in the delta slots, which the
// betweenness pass clears before use.
__device__ bool angular_phase2(const AngArgs &a, CAngJob &jb, const LaneWs &L, uint32_t epoch, uint32_t n_order, uint32_t &fail) {
  char *w = L.w;
  const uint32_t lane = L.lane;
  for (uint32_t q = 0; q < n_order; ++q) {
    const uint32_t ss = *il<uint32_t>(w, a.o_order, q, lane, a.scap);
    *il<double>(w, a.o_delta, ss * a.tcap, lane, (size_t)a.scap * a.tcap) = (double)q;
    *sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_np) = 0;
    *sr<double>(w, a.o_srec, ss, lane, a.scap, kS_sigma) = 0.0;
  }
  // the source's two states are the first two slots
  *sr<double>(w, a.o_srec, 0, lane, a.scap, kS_sigma) = 1.0;
  *sr<double>(w, a.o_srec, 1, lane, a.scap, kS_sigma) = 1.0;
  const float band = 1.0f + jb.tolerance;
  for (uint32_t pos = 0; pos < n_order; ++pos) {
    const uint32_t uss = *il<uint32_t>(w, a.o_order, pos, lane, a.scap);
    const uint32_t us = *sr<uint32_t>(w, a.o_srec, uss, lane, a.scap, kS_state) & 0x7fffffffu;
    const uint32_t u = us >> 1, entry = us & 1u;
    const float u_route = *sr<float>(w, a.o_srec, uss, lane, a.scap, kS_route);
    const double u_sigma = *sr<double>(w, a.o_srec, uss, lane, a.scap, kS_sigma);
    for (uint32_t j = jb.off[u]; j < jb.off[u + 1]; ++j) {
      const uint4 e = jb.adj[j];
      if (((e.x >> 30) & 1u) != 1u - entry) continue;
      const uint32_t v = e.x & 0x3fffffffu;
      const uint32_t vs = 2 * v + (e.x >> 31);
      const uint64_t sk = L.skey[vs];
      if ((uint32_t)(sk >> 32) != epoch) continue;
      const uint32_t vss = (uint32_t)sk;
      if (!(*sr<uint32_t>(w, a.o_srec, vss, lane, a.scap, kS_state) & 0x80000000u)) continue;  // never settled
      if (*il<double>(w, a.o_delta, vss * a.tcap, lane, (size_t)a.scap * a.tcap) <= (double)pos) continue;  // settled before u
      float cr = u_route + __uint_as_float(e.z);
      cr = cr + __uint_as_float(e.w);
      if (!(cr <= *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_route) * band)) continue;
      uint32_t *np = sr<uint32_t>(w, a.o_srec, vss, lane, a.scap, kS_np);
      uint32_t *preds = reinterpret_cast<uint32_t *>(w + a.o_preds) + ((size_t)lane * a.scap + vss) * a.maxp;
      const uint32_t k = *np;
      bool has = false;
      for (uint32_t q = 0; q < k; ++q)
        if (preds[q] == uss) has = true;
      if (has) continue;
      if (k >= a.maxp) {
        fail |= kFailPreds;
        return false;
      }
      preds[k] = uss;
      *np = k + 1;
      double *sg = sr<double>(w, a.o_srec, vss, lane, a.scap, kS_sigma);
      *sg = *sg + u_sigma;
    }
  }
  return true;
}

// QK 1: bucket queue, 0: 4-ary heap.
template <int QK>
__global__ __launch_bounds__(kBlock) void k_angular(AngArgs a) {
  const uint32_t lane = __lane_id();
  const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;  // global lane
  const uint32_t wid = __builtin_amdgcn_readfirstlane(gl / kWave);
  LaneWs L;
  // lanes >= per_wave only ever take padding items and touch no workspace
  L.skey = reinterpret_cast<uint64_t *>(a.ws + ((size_t)wid * a.per_wave + lane) * a.lane_stride);
  L.nkey = L.skey + 2 * (size_t)a.N;
  L.w = a.wave_ws + (size_t)wid * a.wave_stride;
  L.lane = lane;
  unsigned long long st_reached = 0, st_edges = 0, st_within = 0;
  // sources this lane has finished in the call so far (earlier launches included)
  uint32_t pdone = a.progress ? __hip_atomic_load(&a.progress[gl], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) : 0u;
  char *w = L.w;
  // Items are dealt statically, lane gl taking gl, gl + lanes, ...; each job's
  // items are padded to whole waves, so the 64 items of one wave belong to one
  // job and its fields load once per wave into scalar registers.
  const uint32_t lanes_total = gridDim.x * blockDim.x;
  for (uint32_t it = gl; it < a.n_items; it += lanes_total) {
    uint32_t job = 0, si = it;
    if (a.items) {
      const uint2 x = a.items[it];
      job = __builtin_amdgcn_readfirstlane(x.x);
      si = x.y;
      if (si == kNoSource) continue;  // padding
    }
    CAngJob &jb = ((CAngJob *)a.jobs)[job];
    const uint32_t src = jb.sources[si];
    const float wt = jb.src_wt[si];
    // every item of every launch since the maps were zeroed has its own epoch,
    // so map entries left by any earlier search (any layout) never match
    const uint32_t epoch = a.epoch_base + it + 1;
    uint32_t n_states, n_nodes, n_order, fail = 0;
    bool ok;
    auto q = make_queue<QK>(a, jb, w, lane);
    ok = angular_source(a, jb, L, q, src, epoch, n_states, n_nodes, n_order, fail);
    if (ok && jb.tolerance > kTieEps && (jb.flags & 2u)) ok = angular_phase2(a, jb, L, epoch, n_order, fail);
    if (!ok) {
      const uint32_t q = atomicAdd(a.n_fail, 1u);
      a.fail_list[q] = it;
      atomicOr(a.fail_flags, fail);
      continue;
    }
    st_reached += n_nodes;
    st_edges += n_order;
    const uint32_t T = jb.T, TS = jb.tstride;
    // ---- closeness at targets (centrality.rs:1995-2033), kCloseBatch node
    // records loaded before their atomics: vmcnt counts atomics with the
    // loads, in order, so a record loaded behind them waits for them
    for (uint32_t k0 = 1; k0 < n_nodes; k0 += kCloseBatch) {
      uint4 rec[kCloseBatch];
#pragma unroll
      for (uint32_t i = 0; i < kCloseBatch; ++i)
        rec[i] = k0 + i < n_nodes ? *nr<uint4>(w, a.o_nrec, k0 + i, lane, a.ncap, 0) : make_uint4(src, 0u, 0u, 0u);
#pragma unroll
      for (uint32_t i = 0; i < kCloseBatch; ++i) {
        const uint32_t v = rec[i].x;  // node, best route, its seconds
        if (v == src) continue;
        const float bsd = __uint_as_float(rec[i].y);
        const float bs = __uint_as_float(rec[i].z);
        if (!isfinite(bsd) || !isfinite(bs)) continue;
        double *accv = jb.acc + v;
        const size_t NN = jb.N;
        for (uint32_t t = 0; t < T; ++t) {
          if (bs <= jb.sec_thr[t]) {
            ++st_within;
            if (jb.flags & 4u) atomicAdd(&jb.reach[t], 1u);
            if (jb.flags & 1u) {
              unsafeAtomicAdd(&accv[((size_t)0 * TS + t) * NN], (double)wt);
              const float far_ang = jb.offset + (bsd / jb.unit);
              unsafeAtomicAdd(&accv[((size_t)1 * TS + t) * NN], (double)(far_ang * wt));
              const float harm_ang = 1.0f + (bsd / jb.unit);
              unsafeAtomicAdd(&accv[((size_t)2 * TS + t) * NN], (double)((1.0f / harm_ang) * wt));
            }
          }
        }
      }
    }
    // ---- betweenness (centrality.rs:2037-2096, 823-873): every threshold in one
    // walk over the settled states; each threshold keeps its own seeds and
    // dependencies ([T][scap]) and sees its terms in the reference's order
    if (jb.flags & 2u) {
      // seeds and dependencies are state-major: a state's thresholds side by side
      const uint32_t tc = a.tcap;
      for (uint32_t q = 0; q < n_order; ++q) {
        const uint32_t ss = *il<uint32_t>(w, a.o_order, q, lane, a.scap);
        for (uint32_t t = 0; t < T; ++t) {
          *il<double>(w, a.o_delta, ss * tc + t, lane, (size_t)a.scap * a.tcap) = 0.0;
          *il<double>(w, a.o_seed, ss * tc + t, lane, (size_t)a.scap * a.tcap) = 0.0;
        }
      }
      // seeds: best terminal states split by sigma (best_angular_target_states),
      // kCloseBatch nodes at a time: their records, then their state keys, then
      // the candidate states, all loaded before the batch's seed stores
      for (uint32_t k0 = 1; k0 < n_nodes; k0 += kCloseBatch) {
        uint4 rec[kCloseBatch];
#pragma unroll
        for (uint32_t i = 0; i < kCloseBatch; ++i)
          rec[i] = k0 + i < n_nodes ? *nr<uint4>(w, a.o_nrec, k0 + i, lane, a.ncap, 0) : make_uint4(src, 0u, 0u, 0u);
        uint32_t cand[kCloseBatch][2];
        bool has[kCloseBatch][2];
        double pc[kCloseBatch];
#pragma unroll
        for (uint32_t i = 0; i < kCloseBatch; ++i) {
          const uint32_t v = rec[i].x;  // node, best route, its seconds
          const bool live = v != src && isfinite(__uint_as_float(rec[i].y)) && isfinite(__uint_as_float(rec[i].z));
          pc[i] = 0.0;
#pragma unroll
          for (uint32_t slot = 0; slot < 2; ++slot) {
            const uint64_t sk = live ? L.skey[2 * v + slot] : 0ull;
            has[i][slot] = live && (uint32_t)(sk >> 32) == epoch;
            cand[i][slot] = (uint32_t)sk;
          }
          if (live) pc[i] = jb.eligible[v] ? 0.5 : 1.0;
        }
        // the candidates' sigma, seconds and route, read once for every threshold
        double csg[kCloseBatch][2];
        float csec[kCloseBatch][2], crt[kCloseBatch][2];
#pragma unroll
        for (uint32_t i = 0; i < kCloseBatch; ++i) {
#pragma unroll
          for (uint32_t slot = 0; slot < 2; ++slot) {
            csg[i][slot] = 0.0;
            csec[i][slot] = crt[i][slot] = INFINITY;
            if (has[i][slot]) {
              const uint4 hd = *sr<uint4>(w, a.o_srec, cand[i][slot], lane, a.scap, 0);  // route, secs, state, np
              crt[i][slot] = __uint_as_float(hd.x);
              csec[i][slot] = __uint_as_float(hd.y);
              csg[i][slot] = *sr<double>(w, a.o_srec, cand[i][slot], lane, a.scap, kS_sigma);
            }
          }
        }
#pragma unroll
        for (uint32_t i = 0; i < kCloseBatch; ++i) {
          if (pc[i] == 0.0) continue;
          const float br = __uint_as_float(rec[i].y), bs = __uint_as_float(rec[i].z);
          for (uint32_t t = 0; t < T; ++t) {
            const float sthr = jb.sec_thr[t];
            if (bs > sthr) continue;
            // the candidates in slot order, as the reference lists them
            bool use[2];
            double total = 0.0;
#pragma unroll
            for (uint32_t slot = 0; slot < 2; ++slot) {
              use[slot] = has[i][slot] && csg[i][slot] != 0.0 && !(csec[i][slot] > sthr) &&
                          crt[i][slot] <= br * (1.0f + jb.tolerance);
            }
#pragma unroll
            for (uint32_t slot = 0; slot < 2; ++slot)
              if (use[slot]) total += csg[i][slot];
            if (!use[0] && !use[1]) continue;
            if (total == 0.0) continue;
            // a state's seed words are written only here, for its own node, at
            // most once per threshold, onto cleared words: 0 + x is x
#pragma unroll
            for (uint32_t slot = 0; slot < 2; ++slot)
              if (use[slot])
                *il<double>(w, a.o_seed, cand[i][slot] * tc + t, lane, (size_t)a.scap * a.tcap) =
                    pc[i] * (csg[i][slot] / total);
          }
        }
      }
      // states by route descending, equal routes in settle order (stable sort)
      int hi = (int)n_order - 1;
      while (hi >= 0) {
        // routes by settle position: the group scan reads neighbouring words;
        // the group's last state loads with its route (most groups hold one)
        const float r_hi = *il<float>(w, a.o_orte, hi, lane, a.scap);
        const uint32_t ss_hi = *il<uint32_t>(w, a.o_order, hi, lane, a.scap);
        int lo = hi;
        while (lo > 0 && *il<float>(w, a.o_orte, lo - 1, lane, a.scap) == r_hi) --lo;
        for (int q = lo; q <= hi; ++q) {
          const uint32_t ss = q == hi ? ss_hi : *il<uint32_t>(w, a.o_order, q, lane, a.scap);
          const float secs_ss = *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_secs);
          bool any = false;
          for (uint32_t t = 0; t < T; ++t) any = any || secs_ss <= jb.sec_thr[t];
          if (!any) continue;
          const double sw = *sr<double>(w, a.o_srec, ss, lane, a.scap, kS_sigma);
          if (sw == 0.0) continue;
          const uint32_t k = *sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_np);
          const uint32_t *preds = reinterpret_cast<const uint32_t *>(w + a.o_preds) + ((size_t)lane * a.scap + ss) * a.maxp;
          const uint32_t node = (*sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_state) & 0x7fffffffu) >> 1;
          // the state's seed and dependency for every threshold it lies within
          // (0 elsewhere), then each predecessor once: its sigma, then its
          // dependencies for all thresholds together. Each dependency word still
          // takes its terms in walk order, so the sums are those of a walk per
          // threshold.
          double seed[kMaxT], dep[kMaxT];
          bool live = false;
#pragma unroll
          for (uint32_t t = 0; t < kMaxT; ++t) {
            seed[t] = dep[t] = 0.0;
            if (t < T && secs_ss <= jb.sec_thr[t]) {
              seed[t] = *il<double>(w, a.o_seed, ss * tc + t, lane, (size_t)a.scap * a.tcap);
              dep[t] = seed[t] + *il<double>(w, a.o_delta, ss * tc + t, lane, (size_t)a.scap * a.tcap);
            }
            live = live || dep[t] != 0.0;
          }
          if (!live) continue;
          for (uint32_t p = 0; p < k; ++p) {
            const uint32_t ps = preds[p];
            const double sv = *sr<double>(w, a.o_srec, ps, lane, a.scap, kS_sigma);
            if (sv == 0.0) continue;
            const double ratio = sv / sw;
#pragma unroll
            for (uint32_t t = 0; t < kMaxT; ++t) {
              if (dep[t] == 0.0) continue;
              double *dp = il<double>(w, a.o_delta, ps * tc + t, lane, (size_t)a.scap * a.tcap);
              *dp = *dp + ratio * dep[t];
            }
          }
          if (node == src) continue;
#pragma unroll
          for (uint32_t t = 0; t < kMaxT; ++t) {
            const double credit = dep[t] - seed[t];
            if (dep[t] != 0.0 && credit > 0.0)
              unsafeAtomicAdd(&jb.acc[((size_t)3 * jb.tstride + t) * jb.N + node], credit * (double)wt);
          }
        }
        hi = lo - 1;
      }
    }
    if (a.progress) __hip_atomic_store(&a.progress[gl], ++pdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
  atomicAdd(&a.stats[0], st_reached);
  atomicAdd(&a.stats[1], st_edges);
  atomicAdd(&a.stats[2], st_within);
}

// dijkstra_tree_simplest (centrality.rs:1202-1332) for one source on one lane:
// absolute +-1e-4 bands, per-node view collapsed as the reference does.
__global__ void k_tree_simplest(AngArgs a, uint32_t src) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  CAngJob &jb = ((CAngJob *)a.jobs)[0];
  const float eps = kTieEps;
  LaneWs L;
  L.skey = reinterpret_cast<uint64_t *>(a.ws);
  L.nkey = L.skey + 2 * (size_t)a.N;
  L.w = a.wave_ws;
  L.lane = 0;
  char *w = L.w;
  const uint32_t epoch = a.epochs[0] + 1;
  a.epochs[0] = epoch;
  // state slots: route metric (kS_route), agg seconds (kS_secs), simplest
  // distance (o_delta, an f32 value held exactly in an f64 slot)
  uint32_t n_states = 0, no = 0;
  Heap h{w, a.o_heap, 0, 0, a.hcap};
  a.t_order[no++] = src;
  a.t_vis[src] = 1;
  a.t_simpl[src] = 0.0f;
  a.t_secs[src] = 0.0f;
  for (uint32_t slot = 0; slot < 2; ++slot) {
    const uint32_t s = 2 * src + slot;
    const uint32_t ss = n_states++;
    L.skey[s] = ((uint64_t)epoch << 32) | ss;
    *sr<float>(w, a.o_srec, ss, 0, a.scap, kS_route) = 0.0f;            // route metric
    *sr<float>(w, a.o_srec, ss, 0, a.scap, kS_secs) = 0.0f;             // agg seconds
    *il<double>(w, a.o_delta, ss, 0, (size_t)a.scap * a.tcap) = 0.0;            // simpl dist (f32 stored in f64 slot)
    *sr<uint32_t>(w, a.o_srec, ss, 0, a.scap, kS_state) = s;
    h.push(0.0f, 0.0f, s);
  }
  while (h.n > 0) {
    const uint4 top = h.pop();
    const uint32_t us = top.z;
    const uint32_t uss = (uint32_t)L.skey[us];
    uint32_t *ust = sr<uint32_t>(w, a.o_srec, uss, 0, a.scap, kS_state);
    if (*ust & 0x80000000u) continue;
    *ust = us | 0x80000000u;
    const uint32_t u = us >> 1, entry = us & 1u;
    const float u_secs = *sr<float>(w, a.o_srec, uss, 0, a.scap, kS_secs);
    const float u_simpl = (float)*il<double>(w, a.o_delta, uss, 0, (size_t)a.scap * a.tcap);
    for (uint32_t j = jb.off[u]; j < jb.off[u + 1]; ++j) {
      const uint4 e = jb.adj[j];
      if (((e.x >> 30) & 1u) != 1u - entry) continue;
      const uint32_t v = e.x & 0x3fffffffu;
      const uint32_t vs = 2 * v + (e.x >> 31);
      const float cs = u_secs + __uint_as_float(e.y);
      if (cs > jb.maxsec) continue;
      const float csimpl = u_simpl + __uint_as_float(e.z);
      const float cmetric = csimpl + __uint_as_float(e.w);
      const uint64_t sk = L.skey[vs];
      uint32_t vss;
      float vmetric, vsecs;
      if ((uint32_t)(sk >> 32) == epoch) {
        vss = (uint32_t)sk;
        vmetric = *sr<float>(w, a.o_srec, vss, 0, a.scap, kS_route);
        vsecs = *sr<float>(w, a.o_srec, vss, 0, a.scap, kS_secs);
      } else {
        vss = 0xffffffffu;
        vmetric = INFINITY;
        vsecs = INFINITY;
      }
      const bool improved = cmetric + eps < vmetric;
      const bool tied = fabsf(cmetric - vmetric) <= eps;
      if (improved || (tied && cs < vsecs && us != vs)) {
        if (vss == 0xffffffffu) {
          if (n_states >= a.scap || h.n >= a.hcap) {
            atomicOr(a.fail_flags, kFailCap);
            return;
          }
          vss = n_states++;
          L.skey[vs] = ((uint64_t)epoch << 32) | vss;
          *sr<uint32_t>(w, a.o_srec, vss, 0, a.scap, kS_state) = vs;
        }
        *sr<float>(w, a.o_srec, vss, 0, a.scap, kS_route) = cmetric;
        *il<double>(w, a.o_delta, vss, 0, (size_t)a.scap * a.tcap) = (double)csimpl;
        *sr<float>(w, a.o_srec, vss, 0, a.scap, kS_secs) = cs;
        if (!h.push(cmetric, cs, vs)) {
          atomicOr(a.fail_flags, kFailHeap);
          return;
        }
        // collapsed node view (centrality.rs:1307-1326); t_vis: 0 unreached, 1 reached
        const bool first = a.t_vis[v] == 0;
        if (first) a.t_order[no++] = v;
        const float nsimpl = a.t_simpl[v];
        const bool nimp = csimpl + eps < nsimpl;
        const bool ntied = fabsf(csimpl - nsimpl) <= eps;
        if (first || nimp || (ntied && cs < a.t_secs[v])) {
          a.t_vis[v] = 1;
          a.t_simpl[v] = csimpl;
          a.t_secs[v] = cs;
          a.t_pred[v] = (int32_t)u;
        }
      }
    }
  }
  *a.t_norder = no;
}

// dijkstra_tree_simplest (centrality.rs:1202-1332) for many sources, one per
// lane, as angular data aggregation walks it (data.rs:538-548): per source the
// nodes it reaches in first-reach order (visited_nodes) with the seconds of
// their collapsed node view, written as {node | seconds << 32} to the item's
// list of ncap entries, and their number to counts (kNoList when a capacity ran
// out: the item is listed in fail_list and searched again with more room).
// State records keep route, seconds, state and the simplest distance (in the
// predecessor-count word); node records node, simplest distance and seconds.
// QK 1: bucket queue, 0: 4-ary heap.
template <int QK>
__global__ __launch_bounds__(kBlock) void k_reach_simplest(AngArgs a, const uint32_t *items, uint64_t *lists,
                                                           uint32_t *counts) {
  const uint32_t lane = __lane_id();
  const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t wid = __builtin_amdgcn_readfirstlane(gl / kWave);
  LaneWs L;
  L.skey = reinterpret_cast<uint64_t *>(a.ws + ((size_t)wid * kWave + lane) * a.lane_stride);
  L.nkey = L.skey + 2 * (size_t)a.N;
  L.w = a.wave_ws + (size_t)wid * a.wave_stride;
  L.lane = lane;
  char *w = L.w;
  CAngJob &jb = ((CAngJob *)a.jobs)[0];
  const float eps = kTieEps;
  unsigned long long st_reached = 0, st_states = 0;
  const uint32_t lanes_total = gridDim.x * blockDim.x;
  for (uint32_t it = gl; it < a.n_items; it += lanes_total) {
    const uint32_t src = jb.sources[items ? items[it] : it];
    const uint32_t epoch = a.epoch_base + it + 1;
    uint32_t n_states = 0, n_nodes = 1, fail = 0;
    std::conditional_t<QK == 1, BQueue, Heap> h = make_queue<QK>(a, jb, w, lane);
    L.nkey[src] = ((uint64_t)epoch << 32) | 0u;
    *nr<uint32_t>(w, a.o_nrec, 0, lane, a.ncap, kN_node) = src;
    *nr<float>(w, a.o_nrec, 0, lane, a.ncap, kN_brt) = 0.0f;
    *nr<float>(w, a.o_nrec, 0, lane, a.ncap, kN_bsec) = 0.0f;
    for (uint32_t slot = 0; slot < 2; ++slot) {
      const uint32_t s = 2 * src + slot;
      const uint32_t ss = n_states++;
      L.skey[s] = ((uint64_t)epoch << 32) | ss;
      *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_route) = 0.0f;
      *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_secs) = 0.0f;
      *sr<uint32_t>(w, a.o_srec, ss, lane, a.scap, kS_state) = s;
      *sr<float>(w, a.o_srec, ss, lane, a.scap, kS_np) = 0.0f;
      h.push(0.0f, 0.0f, s, ss);
    }
    while (h.n > 0 && !fail) {
      const uint4 top = h.pop();
      const uint32_t us = top.z, uss = top.w;
      uint32_t *ust = sr<uint32_t>(w, a.o_srec, uss, lane, a.scap, kS_state);
      if (*ust & 0x80000000u) continue;
      *ust = us | 0x80000000u;
      const uint32_t u = us >> 1, entry = us & 1u;
      const float u_secs = *sr<float>(w, a.o_srec, uss, lane, a.scap, kS_secs);
      const float u_simpl = *sr<float>(w, a.o_srec, uss, lane, a.scap, kS_np);
      // the state's row (usable edges, row order), or the node's out-row when it spills
      uint4 trow[kRelax];
#pragma unroll
      for (uint32_t q = 0; q < kRelax; ++q) trow[q] = jb.srow[(size_t)us * kRelax + q];
      const bool from_row = trow[0].x == kRowSpill;
      const uint32_t j0 = from_row ? jb.off[u] : 0u, j1 = from_row ? jb.off[u + 1] : kRelax;
      for (uint32_t j = j0; j < j1; ++j) {
        uint4 e;
        if (from_row) {
          e = jb.adj[j];
          if (((e.x >> 30) & 1u) != 1u - entry) continue;
        } else {
          e = trow[0];
#pragma unroll
          for (uint32_t q = 1; q < kRelax; ++q) e = (j == q) ? trow[q] : e;
          if (e.x == kNoEdge) break;  // padding follows the last usable edge
        }
        const uint32_t v = e.x & 0x3fffffffu;
        const uint32_t vs = 2 * v + (e.x >> 31);
        const float cs = u_secs + __uint_as_float(e.y);
        if (cs > jb.maxsec) continue;
        const float csimpl = u_simpl + __uint_as_float(e.z);
        const float cmetric = csimpl + __uint_as_float(e.w);
        const uint64_t sk = L.skey[vs];
        uint32_t vss = 0xffffffffu;
        float vmetric = INFINITY, vsecs = INFINITY;
        if ((uint32_t)(sk >> 32) == epoch) {
          vss = (uint32_t)sk;
          vmetric = *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_route);
          vsecs = *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_secs);
        }
        const bool improved = cmetric + eps < vmetric;
        const bool tied = fabsf(cmetric - vmetric) <= eps;
        if (!(improved || (tied && cs < vsecs && us != vs))) continue;
        if (vss == 0xffffffffu) {
          if (n_states >= a.scap) {
            fail |= kFailCap;
            break;
          }
          vss = n_states++;
          L.skey[vs] = ((uint64_t)epoch << 32) | vss;
          *sr<uint32_t>(w, a.o_srec, vss, lane, a.scap, kS_state) = vs;
        }
        *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_route) = cmetric;
        *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_secs) = cs;
        *sr<float>(w, a.o_srec, vss, lane, a.scap, kS_np) = csimpl;
        if (!h.push(cmetric, cs, vs, vss)) {
          fail |= kFailHeap;
          break;
        }
        // collapsed node view (centrality.rs:1307-1326); a node's slot is its
        // position in visited_nodes
        const uint64_t nk = L.nkey[v];
        if ((uint32_t)(nk >> 32) != epoch) {
          if (n_nodes >= a.ncap) {
            fail |= kFailCap;
            break;
          }
          const uint32_t ns = n_nodes++;
          L.nkey[v] = ((uint64_t)epoch << 32) | ns;
          *nr<uint32_t>(w, a.o_nrec, ns, lane, a.ncap, kN_node) = v;
          *nr<float>(w, a.o_nrec, ns, lane, a.ncap, kN_brt) = csimpl;
          *nr<float>(w, a.o_nrec, ns, lane, a.ncap, kN_bsec) = cs;
        } else {
          const uint32_t ns = (uint32_t)nk;
          float *nsimpl = nr<float>(w, a.o_nrec, ns, lane, a.ncap, kN_brt);
          float *nsecs = nr<float>(w, a.o_nrec, ns, lane, a.ncap, kN_bsec);
          const bool nimp = csimpl + eps < *nsimpl;
          const bool ntied = fabsf(csimpl - *nsimpl) <= eps;
          if (nimp || (ntied && cs < *nsecs)) {
            *nsimpl = csimpl;
            *nsecs = cs;
          }
        }
      }
    }
    if (fail) {
      counts[it] = kNoList;
      const uint32_t q = atomicAdd(a.n_fail, 1u);
      a.fail_list[q] = it;
      atomicOr(a.fail_flags, fail);
      continue;
    }
    uint64_t *lst = lists + (size_t)it * a.ncap;
    // kCloseBatch node records loaded before their list stores (a load issued
    // behind a store waits for it)
    for (uint32_t k0 = 0; k0 < n_nodes; k0 += kCloseBatch) {
      uint4 rec[kCloseBatch];
#pragma unroll
      for (uint32_t i = 0; i < kCloseBatch; ++i)
        if (k0 + i < n_nodes) rec[i] = *nr<uint4>(w, a.o_nrec, k0 + i, lane, a.ncap, 0);  // node, best, its seconds
#pragma unroll
      for (uint32_t i = 0; i < kCloseBatch; ++i)
        if (k0 + i < n_nodes) lst[k0 + i] = (uint64_t)rec[i].x | ((uint64_t)rec[i].z << 32);
    }
    counts[it] = n_nodes;
    st_reached += n_nodes;
    st_states += n_states;
  }
  atomicAdd(&a.stats[0], st_reached);
  atomicAdd(&a.stats[1], st_states);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct AngPlan {
  uint32_t scap, ncap, hcap, maxp, lanes, tcap;
  uint32_t bq_cap = 0;  // bucket queue entries per bucket (0: the heap)
  size_t o_bqe = 0;
  uint32_t per_wave = kWave;  // lane blocks per wave in the interleaved arrays
  size_t lane_stride, wave_stride;
  size_t o_srec, o_preds, o_order, o_orte, o_delta, o_seed, o_nrec, o_heap;
};

void ang_layout(AngPlan &p, uint32_t N, uint32_t T = 1) {
  p.tcap = std::max<uint32_t>(T, 1);
  p.lane_stride = align256(8ull * 3 * N);  // skey [2N] + nkey [N]
  size_t o = 0;
  auto take = [&](size_t &field, size_t elem_bytes, size_t count) {
    field = o;
    o = align256(o + elem_bytes * count * p.per_wave);
  };
  take(p.o_srec, kSRec, p.scap);
  take(p.o_preds, 4, (size_t)p.scap * p.maxp);
  take(p.o_order, 4, p.scap);
  take(p.o_orte, 4, p.scap);
  take(p.o_delta, 8, (size_t)p.scap * std::max<uint32_t>(T, 1));
  take(p.o_seed, 8, (size_t)p.scap * std::max<uint32_t>(T, 1));
  take(p.o_nrec, kNRec, p.ncap);
  if (p.bq_cap) {
    take(p.o_bqe, 16, (size_t)kBuckets * p.bq_cap);
    p.o_heap = 0;
  } else {
    take(p.o_heap, 16, p.hcap);
  }
  p.wave_stride = o;
}

struct AngWorkspace {
  int device = -1;
  char *lane_ws = nullptr, *wave_ws = nullptr;
  size_t lane_bytes = 0, wave_bytes = 0;
  uint64_t epoch_hi = 0;  // epochs issued since the lane maps were last zeroed
  std::mutex mu;
};

std::mutex g_ang_mu;
std::map<int, std::unique_ptr<AngWorkspace>> g_ang;

AngWorkspace &ang_ws(int device) {
  std::lock_guard<std::mutex> lk(g_ang_mu);
  auto &p = g_ang[device];
  if (!p) {
    p = std::make_unique<AngWorkspace>();
    p->device = device;
  }
  return *p;
}

// Grows the workspace to the plan (lane maps of per_wave lanes per wave,
// interleaved arrays per wave) and returns the launch's epoch base. The maps
// are zeroed only when allocated or when the 32-bit epochs would wrap: a
// launch's epochs all exceed every epoch issued before it, so the maps need no
// clearing when the layout (graph size, lanes per wave) changes between calls.
uint32_t prepare_ang(AngWorkspace &w, const AngPlan &p, uint64_t n_items, hipStream_t stream) {
  const size_t lane_need = p.lane_stride * (size_t)(p.lanes / kWave) * p.per_wave;
  const size_t wave_need = p.wave_stride * (p.lanes / kWave);
  bool zero = false;
  if (lane_need > w.lane_bytes) {
    (void)hipFree(w.lane_ws);
    w.lane_ws = nullptr;
    w.lane_bytes = 0;
    if (hipMalloc(&w.lane_ws, lane_need) != hipSuccess) {
      (void)hipGetLastError();
      w.lane_ws = nullptr;
      fail(CS_ENOMEM, "could not allocate the angular search workspace");
    }
    w.lane_bytes = lane_need;
    zero = true;
  }
  if (wave_need > w.wave_bytes) {
    (void)hipFree(w.wave_ws);
    w.wave_ws = nullptr;
    w.wave_bytes = 0;
    if (hipMalloc(&w.wave_ws, wave_need) != hipSuccess) {
      (void)hipGetLastError();
      w.wave_ws = nullptr;
      fail(CS_ENOMEM, "could not allocate the angular search workspace");
    }
    w.wave_bytes = wave_need;
  }
  if (w.epoch_hi + n_items + 2 >= 0xfffffff0ull) zero = true;
  if (zero) {
    hip_check(hipMemsetAsync(w.lane_ws, 0, w.lane_bytes, stream), "memset");
    w.epoch_hi = 0;
  }
  const uint32_t base = (uint32_t)w.epoch_hi;
  w.epoch_hi += n_items + 1;
  return base;
}

}  // namespace

// Forward adjacency of the dual with endpoint slots, angle and tie-break
// increments (built once per speed; centrality.rs:533-566,638-667).
static uint4 *angular_adjacency(Graph &g, DevEdges &de) {
  if (de.ang_adj) return de.ang_adj;
  std::vector<uint4> adj(g.m);
  float min_es = INFINITY, max_w = 0.0f;
  for (uint32_t v = 0; v < g.n; ++v) {
    for (uint32_t j = g.fwd_off[v]; j < g.fwd_off[v + 1]; ++j) {
      const uint32_t e = g.fwd_eid[j];
      const uint32_t u = g.edst[e];
      const float es = edge_seconds(g, v, u, e, de.speed, false);
      const float tb = kAngularTieBreak * g.elen[e];
      uint32_t eb, ab, tbb;
      std::memcpy(&eb, &es, 4);
      std::memcpy(&ab, &g.eang[e], 4);
      std::memcpy(&tbb, &tb, 4);
      adj[j] = make_uint4(u | ((uint32_t)g.slot_src[e] << 30) | ((uint32_t)g.slot_dst[e] << 31), eb, ab, tbb);
      min_es = std::min(min_es, es);
      max_w = std::max(max_w, g.eang[e] + tb);
    }
  }
  hip_check(hipMalloc(&de.ang_adj, sizeof(uint4) * std::max<size_t>(adj.size(), 1)), "hipMalloc");
  if (!adj.empty())
    hip_check(hipMemcpy(de.ang_adj, adj.data(), sizeof(uint4) * adj.size(), hipMemcpyHostToDevice), "upload");
  de.min_es_ang = min_es;
  de.max_w_ang = max_w;
  return de.ang_adj;
}

// State rows of the dual: for state 2v + entry, the out-edges of v that leave
// through its other endpoint (source slot 1 - entry), in row order, padded with
// kNoEdge; kRowSpill first when more than kRelax of them. A search then loads a
// settled state's edges with its record instead of after its row offsets.
static uint4 *angular_state_rows(Graph &g, DevEdges &de) {
  if (de.ang_srow) return de.ang_srow;
  angular_adjacency(g, de);  // de.ang_adj, read back below
  std::vector<uint4> adj(g.m);
  if (g.m) hip_check(hipMemcpy(adj.data(), de.ang_adj, sizeof(uint4) * g.m, hipMemcpyDeviceToHost), "download");
  const uint4 pad = make_uint4(kNoEdge, 0u, 0u, 0u);
  std::vector<uint4> rows((size_t)2 * g.n * kRelax, pad);
  for (uint32_t v = 0; v < g.n; ++v) {
    for (uint32_t entry = 0; entry < 2; ++entry) {
      uint4 *row = &rows[((size_t)2 * v + entry) * kRelax];
      uint32_t k = 0;
      for (uint32_t j = g.fwd_off[v]; j < g.fwd_off[v + 1]; ++j) {
        if (((adj[j].x >> 30) & 1u) != 1u - entry) continue;
        if (k == kRelax) {
          row[0] = make_uint4(kRowSpill, 0u, 0u, 0u);
          break;
        }
        row[k++] = adj[j];
      }
    }
  }
  hip_check(hipMalloc(&de.ang_srow, sizeof(uint4) * std::max<size_t>(rows.size(), 1)), "hipMalloc");
  if (!rows.empty())
    hip_check(hipMemcpy(de.ang_srow, rows.data(), sizeof(uint4) * rows.size(), hipMemcpyHostToDevice), "upload");
  return de.ang_srow;
}

// Items grouped by job (stable) with each job's run padded to whole waves.
// Items grouped by job, `per_wave` real items to each wave of 64 (the rest
// padding), so that the 64 lanes of a wave always belong to one job.
std::vector<uint2> wave_aligned(const std::vector<uint2> &in, size_t n_jobs, uint32_t per_wave) {
  std::vector<std::vector<uint2>> by(n_jobs);
  for (const uint2 &x : in) by[x.x].push_back(x);
  std::vector<uint2> out;
  for (uint32_t j = 0; j < n_jobs; ++j) {
    for (size_t i = 0; i < by[j].size(); ++i) {
      if (i && i % per_wave == 0)
        while (out.size() % kWave) out.push_back(make_uint2(j, kNoSource));
      out.push_back(by[j][i]);
    }
    while (out.size() % kWave) out.push_back(make_uint2(j, kNoSource));
  }
  return out;
}

// Sources per wave. A lane's search is a chain of dependent gathers, and one
// wave-wide gather costs as many cache-line requests as it has active lanes.
// A launch with few sources therefore spreads them over more waves with fewer
// lanes each, so that every SIMD takes part; a full launch keeps 64.
uint32_t sources_per_wave(uint32_t n_real, uint32_t simds) {
  const uint32_t forced = default_tuning().ang_lanes;
  if (forced) return std::min<uint32_t>(forced, kWave);
  const uint32_t waves = std::max<uint32_t>(simds, 1);
  const uint32_t per = (n_real + waves - 1) / waves;
  return std::max<uint32_t>(1, std::min<uint32_t>(kWave, per));
}

void launch_simplest_jobs(const std::vector<AngSpec> &specs, const std::vector<uint2> *explicit_items,
                          uint32_t n_implicit, hipStream_t stream, cs_stats &st, uint32_t *progress) {
  if (specs.empty()) return;
  const int dev0 = specs[0].dg->device;
  int n_cu = 0;
  hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev0), "attribute");
  const uint32_t n_real = explicit_items ? (uint32_t)explicit_items->size() : n_implicit;
  if (n_real == 0) return;
  const uint32_t per_wave = sources_per_wave(n_real, 4u * (uint32_t)n_cu);
  std::vector<uint2> items;
  if (explicit_items) {
    items = wave_aligned(*explicit_items, specs.size(), per_wave);
  } else if (per_wave < kWave) {
    std::vector<uint2> all(n_implicit);
    for (uint32_t i = 0; i < n_implicit; ++i) all[i] = make_uint2(0u, i);
    items = wave_aligned(all, 1, per_wave);
  }
  const bool have_items = !items.empty();
  uint32_t n_list = have_items ? (uint32_t)items.size() : n_implicit;
  uint32_t N = 1, maxp = 1, T = 1;
  std::vector<AngJob> jobs(specs.size());
  for (size_t i = 0; i < specs.size(); ++i) {
    const AngSpec &sp = specs[i];
    Graph &g = *sp.g;
    if (g.n >= (1u << 30)) fail(CS_EINVAL, "dual graph too large for the angular search (node ids need 30 bits)");
    N = std::max(N, g.n);
    maxp = std::max<uint32_t>(maxp, g.max_in_deg);
    T = std::max(T, sp.cs.T);
    AngJob &j = jobs[i];
    std::memset(&j, 0, sizeof(j));
    j.N = g.n;
    j.off = sp.dg->fwd_off;
    j.adj = angular_adjacency(g, *sp.de);
    j.srow = angular_state_rows(g, *sp.de);
    // buckets of 1 / kBucketSpan of the largest route increment: a pushed key
    // lies at most kBucketSpan + 1 buckets (with rounding) past the last pop
    const float mw = sp.de->max_w_ang;
    j.bq_inv = (mw > 0.0f && std::isfinite(mw)) ? (float)kBucketSpan / mw : 0.0f;
    j.eligible = sp.d_eligible;
    j.T = sp.cs.T;
    j.maxsec = sp.cs.maxsec;
    j.tolerance = sp.cs.tolerance;
    j.unit = sp.cs.ang_unit;
    j.offset = sp.cs.far_offset;
    for (uint32_t t = 0; t < kMaxT; ++t) j.sec_thr[t] = t < sp.cs.T ? (float)sp.cs.seconds[t] : 0.0f;
    j.sources = sp.d_sources;
    j.src_wt = sp.d_wt;
    j.flags = (sp.cs.closeness ? 1u : 0u) | (sp.cs.betweenness ? 2u : 0u) | (sp.cs.track_reach ? 4u : 0u);
    j.acc = sp.d_acc;
    j.tstride = sp.tstride ? sp.tstride : sp.cs.T;
    j.reach = sp.d_reach;
  }
  AngPlan p;
  p.maxp = maxp;
  const uint32_t cap0 = default_tuning().cap;  // search states per lane before a retry
  p.scap = std::min<uint32_t>(2 * N, cap0 ? std::max<uint32_t>(cap0, 2) : 8192u);
  p.ncap = std::min<uint32_t>(N, p.scap);
  p.hcap = 2 * p.scap;
  // sources search with the bucket queue first; one that overfills a bucket
  // (or whose graph's costs leave the ring) runs again with the heap
  p.bq_cap = bucket_capacity();
  const int dev = specs[0].dg->device;
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "props");
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "meminfo");
  AngWorkspace &wsp = ang_ws(dev);
  std::lock_guard<std::mutex> lk(wsp.mu);
  uint32_t *d_ctl = nullptr, *d_fail = nullptr;
  uint2 *d_items = nullptr;
  AngJob *d_jobs = nullptr;
  unsigned long long *d_stats = nullptr;
  hip_check(hipMalloc(&d_ctl, 16), "hipMalloc");
  hip_check(hipMalloc(&d_fail, 4ull * n_list), "hipMalloc");
  hip_check(hipMalloc(&d_stats, 32), "hipMalloc");
  hip_check(hipMalloc(&d_jobs, sizeof(AngJob) * jobs.size()), "hipMalloc");
  hip_check(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(AngJob) * jobs.size(), hipMemcpyHostToDevice, stream), "copy");
  if (have_items) {
    hip_check(hipMalloc(&d_items, sizeof(uint2) * n_list), "hipMalloc");
    hip_check(hipMemcpyAsync(d_items, items.data(), sizeof(uint2) * n_list, hipMemcpyHostToDevice, stream), "copy");
  }
  hip_check(hipMemsetAsync(d_stats, 0, 32, stream), "memset");
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  double kms = 0.0;
  for (int attempt = 0; attempt < 6; ++attempt) {
    p.per_wave = per_wave;
    ang_layout(p, N, T);
    // a wave holds per_wave lanes' maps and interleaved arrays
    const size_t wave_bytes = p.lane_stride * p.per_wave + p.wave_stride;
    const size_t budget = free_b / 4 + wsp.lane_bytes + wsp.wave_bytes;
    const size_t wpb = kBlock / kWave;
    size_t waves = std::min<size_t>(budget / wave_bytes, ((size_t)n_list + kWave - 1) / kWave);
    waves = std::min<size_t>(waves, (size_t)prop.multiProcessorCount * 4 * 8);
    waves = (waves + wpb - 1) / wpb * wpb;
    if (waves == 0 || waves * wave_bytes > budget + wave_bytes * wpb)
      fail(CS_ENOMEM, "graph too large for the angular search workspace on this device");
    const uint32_t lanes = (uint32_t)(waves * kWave);
    p.lanes = lanes;
    const uint32_t epoch_base = prepare_ang(wsp, p, n_list, stream);
    hip_check(hipMemsetAsync(d_ctl, 0, 16, stream), "memset");
    AngArgs a;
    std::memset(&a, 0, sizeof(a));
    a.N = N;
    a.jobs = d_jobs;
    a.items = d_items;
    a.n_items = n_list;
    a.stats = d_stats;
    a.ws = wsp.lane_ws;
    a.wave_ws = wsp.wave_ws;
    a.lane_stride = p.lane_stride;
    a.wave_stride = p.wave_stride;
    a.scap = p.scap;
    a.ncap = p.ncap;
    a.hcap = p.hcap;
    a.maxp = p.maxp;
    a.tcap = p.tcap;
    a.o_srec = p.o_srec;
    a.o_preds = p.o_preds;
    a.o_order = p.o_order;
    a.o_orte = p.o_orte;
    a.o_delta = p.o_delta;
    a.o_seed = p.o_seed;
    a.o_nrec = p.o_nrec;
    a.o_heap = p.o_heap;
    a.o_bqe = p.o_bqe;
    a.bq_cap = p.bq_cap;
    a.per_wave = p.per_wave;
    a.epoch_base = epoch_base;
    a.next_source = d_ctl;
    a.n_fail = d_ctl + 1;
    a.fail_flags = d_ctl + 2;
    a.fail_list = d_fail;
    a.progress = lanes <= kProgressSlots ? progress : nullptr;
    hip_check(hipEventRecord(e0, stream), "event");
    if (p.bq_cap)
      hipLaunchKernelGGL(k_angular<1>, dim3(lanes / kBlock), dim3(kBlock), 0, stream, a);
    else
      hipLaunchKernelGGL(k_angular<0>, dim3(lanes / kBlock), dim3(kBlock), 0, stream, a);
    hip_check(hipGetLastError(), "k_angular launch");
    hip_check(hipEventRecord(e1, stream), "event");
    uint32_t ctl[4];
    hip_check(hipMemcpyAsync(ctl, d_ctl, 16, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    kms += ms;
    st.waves = lanes / kWave;
    const uint32_t n_failed = ctl[1], flags = ctl[2];
    if (n_failed == 0) break;
    if (flags & kFailPreds) fail(CS_EINTERNAL, "predecessor list overflow in the angular search");
    if (attempt == 5) fail(CS_EINTERNAL, "angular workspace retries exhausted");
    st.retries += n_failed;
    // retry the failed items with a larger workspace
    std::vector<uint32_t> fl(n_failed);
    hip_check(hipMemcpy(fl.data(), d_fail, 4ull * n_failed, hipMemcpyDeviceToHost), "copy");
    std::vector<uint2> retry(n_failed);
    for (uint32_t i = 0; i < n_failed; ++i) retry[i] = d_items ? items[fl[i]] : make_uint2(0u, fl[i]);
    items = wave_aligned(retry, specs.size(), per_wave);
    (void)hipFree(d_items);
    hip_check(hipMalloc(&d_items, sizeof(uint2) * items.size()), "hipMalloc");
    hip_check(hipMemcpy(d_items, items.data(), sizeof(uint2) * items.size(), hipMemcpyHostToDevice), "copy");
    n_list = (uint32_t)items.size();
    if (flags & kFailCap) p.scap = std::min<uint32_t>(2 * N, p.scap * 4);
    p.ncap = std::min<uint32_t>(N, p.scap);
    if (p.bq_cap && (flags & kFailHeap)) {
      p.bq_cap = 0;  // the heap from here on, at its first size
    } else if (flags & (kFailHeap | kFailCap)) {
      p.hcap = std::max<uint32_t>(p.hcap * 4, 2 * p.scap);
    }
  }
  unsigned long long stv[4];
  hip_check(hipMemcpy(stv, d_stats, 32, hipMemcpyDeviceToHost), "copy");
  st.kernel_ms += kms;
  st.nodes_reached += stv[0];
  st.edges_scanned += stv[1];
  st.within_sum += stv[2];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(d_ctl);
  (void)hipFree(d_fail);
  (void)hipFree(d_stats);
  (void)hipFree(d_jobs);
  (void)hipFree(d_items);
}

void launch_simplest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs, const uint32_t *d_sources,
                     const float *d_wt, uint32_t n_sources, const uint8_t *d_eligible, double *d_acc,
                     uint32_t *d_reach, hipStream_t stream, cs_stats &st, uint32_t tstride, uint32_t *progress) {
  if (n_sources == 0) return;
  std::vector<AngSpec> specs{AngSpec{&g, &dg, &de, cs, d_sources, d_wt, d_eligible, d_acc, d_reach, tstride}};
  launch_simplest_jobs(specs, nullptr, n_sources, stream, st, progress);
}

void launch_tree_simplest(Graph &g, DevGraph &dg, DevEdges &de, uint32_t src, float maxsec, float *d_simpl,
                          float *d_secs, int32_t *d_pred, uint8_t *d_vis, uint32_t *d_order, uint32_t *d_norder,
                          hipStream_t stream) {
  const uint32_t N = g.n;
  uint4 *adj = angular_adjacency(g, de);
  AngPlan p;
  p.maxp = 1;
  p.scap = 2 * N;
  p.ncap = 1;
  p.hcap = 8 * N + 16;
  p.lanes = kWave;
  ang_layout(p, N);
  // single lane: use a private workspace
  char *lane_ws = nullptr, *wave_ws = nullptr;
  uint32_t *epochs = nullptr, *ctl = nullptr;
  hip_check(hipMalloc(&lane_ws, p.lane_stride), "hipMalloc");
  hip_check(hipMalloc(&wave_ws, p.wave_stride), "hipMalloc");
  hip_check(hipMalloc(&epochs, 4), "hipMalloc");
  hip_check(hipMalloc(&ctl, 16), "hipMalloc");
  hip_check(hipMemsetAsync(lane_ws, 0, p.lane_stride, stream), "memset");
  hip_check(hipMemsetAsync(epochs, 0, 4, stream), "memset");
  hip_check(hipMemsetAsync(ctl, 0, 16, stream), "memset");
  AngJob job;
  std::memset(&job, 0, sizeof(job));
  job.N = N;
  job.off = dg.fwd_off;
  job.adj = adj;
  job.maxsec = maxsec;
  AngJob *d_job = nullptr;
  hip_check(hipMalloc(&d_job, sizeof(AngJob)), "hipMalloc");
  hip_check(hipMemcpyAsync(d_job, &job, sizeof(AngJob), hipMemcpyHostToDevice, stream), "copy");
  AngArgs a;
  std::memset(&a, 0, sizeof(a));
  a.N = N;
  a.jobs = d_job;
  a.ws = lane_ws;
  a.wave_ws = wave_ws;
  a.lane_stride = p.lane_stride;
  a.wave_stride = p.wave_stride;
  a.scap = p.scap;
  a.ncap = p.ncap;
  a.hcap = p.hcap;
  a.maxp = p.maxp;
  a.tcap = p.tcap;
  a.o_srec = p.o_srec;
  a.o_preds = p.o_preds;
  a.o_order = p.o_order;
  a.o_orte = p.o_orte;
  a.o_delta = p.o_delta;
  a.o_seed = p.o_seed;
  a.o_nrec = p.o_nrec;
  a.o_heap = p.o_heap;
  a.epochs = epochs;
  a.fail_flags = ctl + 2;
  a.t_simpl = d_simpl;
  a.t_secs = d_secs;
  a.t_pred = d_pred;
  a.t_vis = d_vis;
  a.t_order = d_order;
  a.t_norder = d_norder;
  hipLaunchKernelGGL(k_tree_simplest, dim3(1), dim3(kWave), 0, stream, a, src);
  hip_check(hipGetLastError(), "k_tree_simplest launch");
  uint32_t hctl[4];
  hip_check(hipMemcpyAsync(hctl, ctl, 16, hipMemcpyDeviceToHost, stream), "copy");
  hip_check(hipStreamSynchronize(stream), "sync");
  (void)hipFree(lane_ws);
  (void)hipFree(wave_ws);
  (void)hipFree(epochs);
  (void)hipFree(ctl);
  (void)hipFree(d_job);
  if (hctl[2]) fail(CS_EINTERNAL, "angular tree workspace exceeded");
}

static_assert(kFailCap == kReachFailCap && kFailHeap == kReachFailHeap, "reach failure bits");

std::vector<uint32_t> launch_reach_simplest(Graph &g, DevGraph &dg, DevEdges &de, float maxsec,
                                            const uint32_t *d_src, const uint32_t *d_items, uint32_t n_items,
                                            uint32_t cap, uint32_t hcap, bool heap, uint64_t *d_lists,
                                            uint32_t *d_counts, hipStream_t stream, cs_stats &st,
                                            uint32_t *fail_bits) {
  *fail_bits = 0;
  if (n_items == 0) return {};
  const uint32_t N = g.n;
  if (N >= (1u << 30)) fail(CS_EINVAL, "dual graph too large for the angular search (node ids need 30 bits)");
  const int dev = dg.device;
  AngJob job;
  std::memset(&job, 0, sizeof(job));
  job.N = N;
  job.off = dg.fwd_off;
  job.adj = angular_adjacency(g, de);
  job.maxsec = maxsec;
  job.sources = d_src;
  job.srow = angular_state_rows(g, de);
  const float mw = de.max_w_ang;
  job.bq_inv = (mw > 0.0f && std::isfinite(mw)) ? (float)kBucketSpan / mw : 0.0f;
  // state records, node records and the queue (bucket ring or heap) of each lane, interleaved per wave
  AngPlan p;
  p.per_wave = kWave;
  p.scap = std::min<uint32_t>(2 * N, 2 * cap);
  p.ncap = cap;
  p.hcap = heap ? hcap : 0;
  p.bq_cap = heap ? 0 : bucket_capacity();
  p.lane_stride = align256(8ull * 3 * N);
  p.o_srec = 0;
  p.o_nrec = align256((size_t)kSRec * p.scap * kWave);
  const size_t o_q = p.o_nrec + align256((size_t)kNRec * p.ncap * kWave);
  if (heap) {
    p.o_heap = o_q;
    p.wave_stride = o_q + align256((size_t)16 * p.hcap * kWave);
  } else {
    p.o_bqe = o_q;
    p.wave_stride = o_q + align256((size_t)16 * kBuckets * p.bq_cap * kWave);
  }
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "props");
  size_t free_b = 0, total_b = 0;
  hip_check(hipMemGetInfo(&free_b, &total_b), "meminfo");
  AngWorkspace &wsp = ang_ws(dev);
  std::lock_guard<std::mutex> lk(wsp.mu);
  const size_t wave_bytes = p.lane_stride * kWave + p.wave_stride;
  const size_t budget = free_b / 4 + wsp.lane_bytes + wsp.wave_bytes;
  const size_t wpb = kBlock / kWave;
  size_t waves = std::min<size_t>(budget / wave_bytes, ((size_t)n_items + kWave - 1) / kWave);
  waves = std::min<size_t>(waves, (size_t)prop.multiProcessorCount * 4 * 8);
  waves = (waves + wpb - 1) / wpb * wpb;
  if (waves == 0 || waves * wave_bytes > budget + wave_bytes * wpb)
    fail(CS_ENOMEM, "graph too large for the angular search workspace on this device");
  p.lanes = (uint32_t)(waves * kWave);
  const uint32_t epoch_base = prepare_ang(wsp, p, n_items, stream);
  DevMem<uint32_t> d_ctl(4), d_fail(n_items);
  DevMem<unsigned long long> d_stats(4);
  DevMem<AngJob> d_job(1);
  hip_check(hipMemsetAsync(d_ctl.p, 0, 16, stream), "memset");
  hip_check(hipMemsetAsync(d_stats.p, 0, 32, stream), "memset");
  hip_check(hipMemcpyAsync(d_job.p, &job, sizeof(AngJob), hipMemcpyHostToDevice, stream), "copy");
  AngArgs a;
  std::memset(&a, 0, sizeof(a));
  a.N = N;
  a.jobs = d_job.p;
  a.n_items = n_items;
  a.stats = d_stats.p;
  a.ws = wsp.lane_ws;
  a.wave_ws = wsp.wave_ws;
  a.lane_stride = p.lane_stride;
  a.wave_stride = p.wave_stride;
  a.scap = p.scap;
  a.ncap = p.ncap;
  a.hcap = p.hcap;
  a.o_srec = p.o_srec;
  a.o_nrec = p.o_nrec;
  a.o_heap = p.o_heap;
  a.o_bqe = p.o_bqe;
  a.bq_cap = p.bq_cap;
  a.per_wave = kWave;
  a.epoch_base = epoch_base;
  a.n_fail = d_ctl.p + 1;
  a.fail_flags = d_ctl.p + 2;
  a.fail_list = d_fail.p;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  hip_check(hipEventRecord(e0, stream), "event");
  if (heap)
    hipLaunchKernelGGL(k_reach_simplest<0>, dim3(p.lanes / kBlock), dim3(kBlock), 0, stream, a, d_items, d_lists,
                       d_counts);
  else
    hipLaunchKernelGGL(k_reach_simplest<1>, dim3(p.lanes / kBlock), dim3(kBlock), 0, stream, a, d_items, d_lists,
                       d_counts);
  hip_check(hipGetLastError(), "k_reach_simplest launch");
  hip_check(hipEventRecord(e1, stream), "event");
  uint32_t ctl[4];
  unsigned long long stv[4];
  hip_check(hipMemcpyAsync(ctl, d_ctl.p, 16, hipMemcpyDeviceToHost, stream), "copy");
  hip_check(hipMemcpyAsync(stv, d_stats.p, 32, hipMemcpyDeviceToHost, stream), "copy");
  hip_check(hipStreamSynchronize(stream), "sync");
  float ms = 0.0f;
  hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  st.kernel_ms += ms;
  st.nodes_reached += stv[0];
  st.edges_scanned += stv[1];
  st.waves = std::max<uint64_t>(st.waves, waves);
  *fail_bits = ctl[2];
  std::vector<uint32_t> failed(ctl[1]);
  if (ctl[1]) hip_check(hipMemcpy(failed.data(), d_fail.p, 4ull * ctl[1], hipMemcpyDeviceToHost), "copy");
  return failed;
}

void release_angular_workspaces() {
  std::lock_guard<std::mutex> lk(g_ang_mu);
  for (auto &dv : g_ang) {
    AngWorkspace &w = *dv.second;
    std::lock_guard<std::mutex> wl(w.mu);
    if (w.lane_ws) (void)hipFree(w.lane_ws);
    if (w.wave_ws) (void)hipFree(w.wave_ws);
    w.lane_ws = w.wave_ws = nullptr;
    w.lane_bytes = w.wave_bytes = 0;
    w.epoch_hi = 0;
  }
}

}  // namespace cs
