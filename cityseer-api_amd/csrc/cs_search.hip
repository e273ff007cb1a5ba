// cs_search.hip -- distance-bounded per-source search on gfx950.
//
// One wavefront (64 lanes) owns one source at a time and pulls the next source
// from a counter of its XCD's share of the source list when it finishes
// (persistent waves, dynamic balance, work stealing across XCDs at the end).
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
//  * Per-wave node state is one epoch-tagged 64-bit key per node (epoch,
//    settled bit, seconds, slot, tie mark) updated with a 64-bit atomic max:
//    newer epochs, settled keys and smaller seconds win, zero-filled memory is
//    stale, and no O(N) state is cleared per source. Settled nodes get compact
//    slots in settle order; the source is always slot 0 (it is alone in bucket
//    0). An offer names the slot that makes it, so a node without a tie mark
//    takes its one predecessor from its key, without looking at its out-row.
//  * Relaxations run edge-parallel over the lanes of a batch (one in-edge per
//    lane); a relaxation loads the neighbour's key first and runs the atomic
//    max only for an offer that may win. Bucket ring counters live in LDS.
//  * Closeness terms are scattered to targets with f64 atomics
//    (centrality.rs:1755-1778); betweenness back-propagates bucket by bucket
//    in reverse (centrality.rs:823-873), reading the predecessor's slot that
//    the forward fold stored.
#include "cs_internal.hpp"

#include <algorithm>

// 64-bit wave reductions of the device library (declared by HIP only under an
// opt-in macro).
extern "C" __device__ uint32_t __ockl_wfscan_add_u32(uint32_t, bool);
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(unsigned long long);
extern "C" __device__ __attribute__((const)) double __ockl_wfred_add_f64(double);
#include <cmath>
#include <cstring>

namespace cs {

namespace {

constexpr uint32_t kFailCap = 1u, kFailBucket = 2u, kFailRing = 4u, kFailInternal = 8u, kFailWide = 16u;
constexpr int kChunk = 1;           // in-edges per lane per step of the segment closeness walk
constexpr uint32_t kXcd = 8;   // XCDs; workgroups are dealt to them round-robin
constexpr uint32_t kScanQ = 4;     // key words per lane per step of the node-ordered result flush
constexpr uint32_t kFlushList = kScanQ * 64;  // LDS list of one flush step (reached nodes, index order)
constexpr uint32_t kFlushWords = 3 * kFlushList;  // {node, seconds bits, slot} per list entry
constexpr uint32_t kRelQ = 3;       // in-edges per lane per relaxation step
constexpr uint32_t kRelWords = 6 * 64;  // per lane {first edge, in-degree, node, seconds, row start, slot} of a batch
constexpr uint32_t kWaveLds = kFlushWords + kRelWords;  // LDS words per wave besides the ring
constexpr uint32_t kLdsWords = 16384;                     // LDS words one workgroup may hold (64 KiB)
constexpr uint32_t kScanSpan = 16; // scan the index span only while it is at most this many times the reach
constexpr uint32_t kEpochMax = 511u;        // 9-bit epoch tags 1..511; 0 is never issued
constexpr uint32_t kEpochShift = 55;        // epoch position in node keys and group records
constexpr size_t kRecBytes = 8;             // one node record: its key
constexpr uint32_t kSlotLimit = 1u << 22;   // 22-bit slots in the node key

enum Mode { kBrandes = 0, kSegment = 1, kTree = 2, kBrandesExport = 3, kAccess = 4 };
// out-row candidates whose loads one fold step batches: two in the Brandes
// search (room for five waves per SIMD), four in the trees and segment search
template <int MODE>
constexpr uint32_t fold_chunk() { return MODE == kBrandes ? 2u : 4u; }

// One job: a graph and the parameters of one centrality request over it. A
// launch runs the sources of one job (plain calls) or of many (batches).
struct Job {
  const uint4 *rows;  // {fwd begin, fwd end, rev begin, rev end} per node
  const uint2 *rev_adj;
  const uint2 *fwd_adj;
  const float2 *rev_tw;
  const float *fwd_twlen;
  const uint32_t *fwd_eid;
  const uint8_t *eligible;
  const uint32_t *sources;
  const float *src_wt;
  const float *src_cyc;
  double *acc;  // [M][tstride][nacc], metric-major like the reference's result rows
  uint32_t *reach;
  float maxsec, speed, inv_delta;
  float tol;  // tie tolerance; above the internal band, Phase 2 rebuilds predecessors
  uint32_t nbk, maxb;
  uint32_t flags;  // 1 closeness, 2 betweenness, 4 track reach, 8 OD-weighted seeds, 16 fold after the search,
                   // 32 some out-edge counts toward the circuit rank (one-way canonical edges)
  double far_max;  // far_factor of the graph's largest out-degree: the widest tie band of any fold
  uint32_t n_sources;
  // OD-weighted betweenness: source position i's destinations and trip weights
  // are od_dest / od_w[od_off[i] .. od_off[i + 1])
  const uint32_t *od_off;
  const uint32_t *od_dest;
  const float *od_w;
  const uint32_t *orig;  // [N] original index of each device node (exact ties)
  uint32_t tstride;  // threshold stride of acc (>= the job's threshold count)
  uint32_t nacc;     // node count of acc (the job's graph)
  float dist[kMaxT];  // thresholds beyond the job's own count hold -1 (never within)
  float beta[kMaxT];
};

// Jobs are read through the constant address space, like kernel arguments:
// scalar loads the compiler may repeat instead of holding every field.
typedef __attribute__((address_space(4))) const Job CJob;

struct SearchArgs {
  const Job *jobs;
  const uint2 *items;  // {job, source position}; null: job 0, position = item index
  uint32_t n_items;
  uint32_t bcap, cap, nbk_lds;
  unsigned long long *stats;  // reached, edges, within
  uint64_t *keys;   // [waves][ncap] node records (epoch-tagged keys)
  size_t ncap;
  char *ws;         // [waves][ws_stride] scratch
  size_t ws_stride;
  size_t o_rec, o_pm, o_acc, o_ods, o_sig, o_bkt, o_bnd, o_lvl, o_bk;
  size_t o_st;       // land-use statistics: per-wave lists (winner count, winners, sort keys, values, weights)
  uint32_t st_cap;   // winners one source may have in those lists
  uint32_t *epochs;
  uint32_t *next;            // [kXcd] per-XCD item counters
  uint32_t xbeg[kXcd + 1];   // XCD x owns item indices [xbeg[x], xbeg[x+1])
  uint32_t xsplit;           // 0: every wave draws from share 0 (one global counter)
  uint32_t *fail_list;       // failed item indices
  uint32_t *n_fail;
  uint32_t *fail_flags;
  TreeOut tout;
  AccessOut acc;
  uint32_t *progress;  // [waves] sources each wave has finished (pinned host memory), or null
};

// The reference's f32 exponential is the host libm's expf, correctly rounded
// in all but a sliver of cases. The device's f32 exp differs from it by up to
// an ulp, which the differences of exponentials in the segment integrals
// magnify; the f64 exponential rounded to f32 agrees with the correctly
// rounded value except when it lands within 2^-29 of a rounding boundary.
__device__ __forceinline__ float exp_f32(float x) { return (float)exp((double)x); }

// Wave-private workspace: only the owning wave reads or writes it, so
// workgroup scope (coherent in the CU's L1 and the XCD's L2) is sufficient.
__device__ __forceinline__ uint64_t ld64(const uint64_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t ld32(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
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
__device__ __forceinline__ void wadd_f64(double *p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// LDS ring counters (one wave's private row).
__device__ __forceinline__ uint32_t lds_ld(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_st(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ uint32_t lds_add(uint32_t *p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Results shared by all waves (accumulators): device scope.
__device__ __forceinline__ void gadd_f64(double *p, double v) { unsafeAtomicAdd(p, v); }
// Orders this wave's earlier memory operations before its later ones (all lanes).
__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
// Wave-wide reductions over the active lanes (the device library's DPP
// sequences: no LDS round trips), result in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return __ockl_wfred_add_u32(v); }
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) { return __ockl_wfred_add_u64(v); }
__device__ __forceinline__ double wave_sum(double v) { return __ockl_wfred_add_f64(v); }
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t wave_or(uint32_t v) { return __ockl_wfred_or_u32(v); }

// Node key, one 64-bit word: epoch (9 bits) | settled (1) | low 31 bits of the
// complemented f32 seconds (31; seconds are never negative, so the dropped top
// bit is always 1) | slot (22) | tie mark (1). A 64-bit atomic max keeps the
// newest epoch, then a settled key over any offer, then the smallest seconds.
// An offer carries the slot of the node that makes it (the predecessor it
// proposes); a settled key carries the node's own slot. The tie mark (set by an
// atomic or) says that two offers within the widest tie band met at the node.
constexpr uint64_t kSettledBit = 1ull << 54;
constexpr uint64_t kSlotMask = (uint64_t)kSlotLimit - 1u;
constexpr uint64_t kLowMask = (kSlotMask << 1) | 1u;  // slot and tie mark
__device__ __forceinline__ uint64_t offer_key_bits(uint32_t epoch, uint32_t secs_bits, uint32_t slot) {
  return ((uint64_t)epoch << kEpochShift) | ((uint64_t)(~secs_bits & 0x7fffffffu) << 23) | ((uint64_t)slot << 1);
}
__device__ __forceinline__ uint64_t offer_key(uint32_t epoch, float secs, uint32_t slot) {
  return offer_key_bits(epoch, __float_as_uint(secs), slot);
}
// The settled form of an offer key: same epoch and seconds, the node's slot.
__device__ __forceinline__ uint64_t settle_key(uint64_t k, uint32_t slot) {
  return (k & ~kLowMask) | kSettledBit | ((uint64_t)slot << 1);
}
__device__ __forceinline__ bool key_current(uint64_t k, uint32_t epoch) { return (uint32_t)(k >> kEpochShift) == epoch; }
__device__ __forceinline__ bool key_settled(uint64_t k) { return (k & kSettledBit) != 0; }
__device__ __forceinline__ bool key_marked(uint64_t k) { return (k & 1u) != 0; }
__device__ __forceinline__ float key_secs(uint64_t k) { return __uint_as_float(~((uint32_t)(k >> 23) | 0x80000000u)); }
__device__ __forceinline__ uint32_t key_slot(uint64_t k) { return (uint32_t)(k >> 1) & (kSlotLimit - 1u); }

__device__ __forceinline__ uint32_t bucket_of(float s, float inv_delta) {
  return (uint32_t)(s * inv_delta);
}

// Reference settle order: ascending seconds (route is monotone in seconds),
// then node index (deterministic stand-in for BinaryHeap's internal order).
// Node ids here are device ids; exact ties compare the original indices.
__device__ __forceinline__ bool key_less(const uint32_t *orig, float sa, uint32_t a, float sb, uint32_t b) {
  return sa < sb || (sa == sb && orig[a] < orig[b]);
}

// Per-slot record, two 64-bit words:
//   lo = node | secs bits << 32
//   hi = Brandes: first predecessor slot | (n_preds | eligible << 8) << 32
//        trees:   predecessor slot | row position of the tree edge << 32
struct WaveWs {
  uint64_t *node;   // [N] node keys (records after the graph's nodes: dedupe groups)
  uint64_t *rec;    // [cap][2]
  uint64_t *pm;     // [cap] predecessor row mask (Brandes) / {origin len, last len} (trees)
  double *acc;      // Brandes: [2T][cap] dependencies, plain then beta-weighted per threshold;
                    // segment: [cap][T] subtree sums
  uint32_t dcap;    // slots per dependency row (cap)
  double *ods;      // [cap] OD trip weight seeding each slot (OD-weighted betweenness)
  double *sig;      // [cap] sigma of each slot (Brandes searches)
  uint2 *bkt;       // [nbk][bcap] bucket ring entries {node, secs bits}
  uint32_t *bnd;    // [maxb + 2] slot bounds per opened bucket
  uint32_t *lvl;    // [maxb + 2] highest Phase 2 resolution round per bucket
  uint32_t *bk;     // [maxb + 2] absolute bucket number of each opened bucket
  uint32_t *ring;   // LDS [nbk] bucket ring counters
  uint32_t *list;   // LDS [3][kFlushList] {node, seconds bits, slot} list of the result flush
  uint32_t *rel;    // LDS [5][64] edge-parallel relaxation table of one batch
};

// Brandes dependency k of slot s: k = t (plain) or T + t (beta-weighted). Rows
// are threshold-major: slots are in bucket order, so the slots within a
// threshold fill its rows up to the bucket straddling it, and the words a
// source touches are dense.
__device__ __forceinline__ double *dep_at(const WaveWs &w, uint32_t k, uint32_t s) {
  return &w.acc[(size_t)k * w.dcap + s];
}

__device__ __forceinline__ WaveWs wave_ws(const SearchArgs &a, uint32_t wid, uint32_t *ring) {
  char *base = a.ws + (size_t)wid * a.ws_stride;
  WaveWs w;
  w.node = a.keys + (size_t)wid * a.ncap;
  w.rec = reinterpret_cast<uint64_t *>(base + a.o_rec);
  w.pm = reinterpret_cast<uint64_t *>(base + a.o_pm);
  w.acc = reinterpret_cast<double *>(base + a.o_acc);
  w.dcap = a.cap;
  w.ods = reinterpret_cast<double *>(base + a.o_ods);
  w.sig = reinterpret_cast<double *>(base + a.o_sig);
  w.bkt = reinterpret_cast<uint2 *>(base + a.o_bkt);
  w.bnd = reinterpret_cast<uint32_t *>(base + a.o_bnd);
  w.lvl = reinterpret_cast<uint32_t *>(base + a.o_lvl);
  w.bk = reinterpret_cast<uint32_t *>(base + a.o_bk);
  w.ring = ring;
  w.list = ring + a.nbk_lds;
  w.rel = w.list + kFlushWords;
  return w;
}

// Predecessor set of a settled node, over the positions of its out-row: a
// bitmask when the row has at most 64 entries; for a wider row, up to four
// positions of 16 bits each (position + 1, packed from the low end in the
// order they joined, 0 after the last). A fifth predecessor at a wide row is
// refused (kFailWide).
constexpr uint32_t kWideRow = 64u, kWideSlots = 4u;  // rows up to 65,534 entries (checked on the host)
__device__ __forceinline__ uint64_t ps_single(uint32_t p, bool wide) {
  return wide ? (uint64_t)(p + 1u) : 1ull << p;
}
__device__ __forceinline__ uint64_t ps_add(uint64_t m, uint32_t p, bool wide, bool &overflow) {
  if (!wide) return m | (1ull << p);
  for (uint32_t k = 0; k < kWideSlots; ++k) {
    const uint32_t x = (uint32_t)(m >> (16u * k)) & 0xffffu;
    if (x == p + 1u) return m;
    if (x == 0u) return m | ((uint64_t)(p + 1u) << (16u * k));
  }
  overflow = true;
  return m;
}
__device__ __forceinline__ uint32_t ps_count(uint64_t m, bool wide) {
  if (!wide) return (uint32_t)__popcll(m);
  uint32_t n = 0;
  for (; m; m >>= 16) ++n;
  return n;
}
__device__ __forceinline__ uint32_t ps_first(uint64_t m, bool wide) {
  return wide ? (uint32_t)(m & 0xffffu) - 1u : (uint32_t)__builtin_ctzll(m);
}
// Takes the next position out of m; false once m is empty.
__device__ __forceinline__ bool ps_next(uint64_t &m, bool wide, uint32_t &p) {
  if (!m) return false;
  if (wide) {
    p = (uint32_t)(m & 0xffffu) - 1u;
    m >>= 16;
  } else {
    p = (uint32_t)__builtin_ctzll(m);
    m &= m - 1;
  }
  return true;
}

// Multi-predecessor fold of one candidate, exactly as the reference's Phase 1
// relaxation of node nb from settled node u (centrality.rs:1413-1437).
struct Fold {
  float best;
  double sigma;
  uint64_t pmask;
  bool wide, overflow;
};

__device__ __forceinline__ void fold_apply(Fold &f, float cs, double sigma_u, uint32_t j,
                                           bool u_in_preds) {
  const float one_m = 1.0f - kTieEps;
  const float one_p = 1.0f + kTieEps;
  if (cs < f.best) {
    if (cs < f.best * one_m) {
      f.pmask = ps_single(j, f.wide);
      f.sigma = sigma_u;
    } else {
      f.pmask = ps_add(f.pmask, j, f.wide, f.overflow);
      f.sigma += sigma_u;
    }
    f.best = cs;
  } else if (cs <= f.best * one_p && !u_in_preds) {
    f.pmask = ps_add(f.pmask, j, f.wide, f.overflow);
    f.sigma += sigma_u;
  }
}

// A candidate offering more than final * (1+1e-4)^(deg+3) can neither stay a
// predecessor nor change which candidates do: every later reset that must
// remove it also removes anything it influenced. Only "near" candidates need
// the ordered fold; usually there is exactly one.
__device__ __forceinline__ double far_factor(uint32_t deg) {
  double f = 1.0;
  for (uint32_t i = 0; i < deg + 3; ++i) f *= 1.00010002;
  return f;
}

// Ordered fold over the near candidates of v (two or more): selection passes in
// settle order (secs_u, u), ties of one u (parallel edges) in row order.
__device__ __forceinline__ uint64_t *nkey(const WaveWs &w, uint32_t v) { return &w.node[v]; }
// Origin-segment length of a settled slot (trees: the high word of its pm entry).
__device__ __forceinline__ float slot_olen(const WaveWs &w, uint32_t s) {
  return __uint_as_float((uint32_t)(ld64(&w.pm[s]) >> 32));
}

// Candidates are out-row nodes settled before v: in an earlier bucket than kb,
// or, with by_key, anywhere before v in (seconds, original index) order.
__device__ bool brandes_fold_exact(CJob &j, const WaveWs &w, uint32_t tag, uint32_t v, float secs_v,
                                   uint32_t kb, uint32_t r0, uint32_t r1, double thr_far, uint32_t n_near,
                                   Fold &f, bool by_key = false) {
  f.best = INFINITY;
  f.sigma = 0.0;
  f.pmask = 0;
  f.wide = (r1 - r0) > kWideRow;
  f.overflow = false;
  bool have_last = false;
  float lsu = 0.0f;
  uint32_t lu = 0, lj = 0;
  for (uint32_t step = 0; step < n_near; ++step) {
    bool found = false;
    float psu = INFINITY, pcs = 0.0f;
    uint32_t pu = 0, pj = 0, psl = 0;
    for (uint32_t p = r0; p < r1; ++p) {
      const uint2 e = j.fwd_adj[p];
      const uint32_t u = e.x & kNodeMask;
      if (u == v) continue;
      const uint64_t ku = ld64(nkey(w, u));
      if (!key_current(ku, tag)) continue;
      const float su = key_secs(ku);
      if (by_key ? !key_less(j.orig, su, u, secs_v, v) : bucket_of(su, j.inv_delta) >= kb) continue;
      const float cs = su + __uint_as_float(e.y);
      if (cs > j.maxsec || (double)cs > thr_far) continue;
      const uint32_t jr = p - r0;
      if (have_last && !(key_less(j.orig, lsu, lu, su, u) || (su == lsu && u == lu && jr > lj))) continue;
      if (!found || key_less(j.orig, su, u, psu, pu) || (su == psu && u == pu && jr < pj)) {
        found = true;
        psu = su;
        pu = u;
        pj = jr;
        pcs = cs;
        psl = key_slot(ku);
      }
    }
    if (!found) return false;
    bool contains = false;
    uint32_t jb;
    for (uint64_t m = f.pmask; ps_next(m, f.wide, jb);)
      if ((j.fwd_adj[r0 + jb].x & kNodeMask) == pu) contains = true;
    const double sg = ldd(&w.sig[psl]);
    fold_apply(f, pcs, sg, pj, contains);
    have_last = true;
    lsu = psu;
    lu = pu;
    lj = pj;
  }
  return f.best == secs_v && !f.overflow;
}

struct SrcCtx {
  uint32_t src;
  uint32_t tag;
  float wt;        // source weight (closeness terms)
  uint32_t k;      // current absolute bucket
  uint32_t nslot;  // settled nodes so far
  uint32_t nb;     // buckets opened
  uint32_t fail;
  uint32_t vlo, vhi;  // smallest and largest settled node index
};

// ---------------------------------------------------------------------------
// Search one source: bucket by bucket, each entry of the bucket is validated
// (its key still equals the pushed value), given a slot, folded and relaxed in
// one pass. NT = number of thresholds (compile time).
// Returns false if the source must be retried with a larger workspace.
// ---------------------------------------------------------------------------
template <int MODE, int NT>
__device__ bool search_source(const SearchArgs &a, CJob &j, const WaveWs &w, SrcCtx &c, uint32_t (&nband)[NT],
                              uint32_t (&eband)[NT], uint32_t &st_edges) {
  constexpr bool kMulti = (MODE == kBrandes || MODE == kBrandesExport);
  const uint32_t lane = lane_id();
  const float speed = j.speed;
  const bool cycles = (MODE == kBrandes) && (j.flags & 1u);
  // with Phase 2 ahead, the Phase 1 fold's predecessors would be discarded
  // Phase 2, or Phase 1 folded after the search (flag 16): no fold at settle time
  const bool phase2 = kMulti && (j.tol > kTieEps || (j.flags & 16u));
  // land-use walks read only each node's seconds: no predecessor at all
  const bool fold = ((MODE != kBrandes) || (j.flags & 2u)) && !phase2 && MODE != kAccess;
  // Brandes folds at settle time take the predecessor an offer key proposes
  // unless the node carries a tie mark (two offers within the widest tie band)
  const bool hint = (MODE == kBrandes) && fold;
  // out-rows are scanned for the circuit rank only where one-way canonical
  // edges exist (two-way keys are all counted by the relaxations)
  const bool out_canon = cycles && (j.flags & 32u);
  if (lane == 0) {
    st64(nkey(w, c.src), offer_key(c.tag, 0.0f, 0u));
    st64(reinterpret_cast<uint64_t *>(&w.bkt[0]), (uint64_t)c.src);
    lds_st(&w.ring[0], 1u);
  }
  wave_fence();
  uint32_t pending = 1;
  c.k = 0;
  c.nslot = 0;
  c.nb = 0;
  c.fail = 0;
  uint32_t vlo = 0xffffffffu, vhi = 0u;  // this lane's settled index bounds
  while (pending > 0) {
    // every bucket index is below maxsec / delta: a larger k means a broken count
    if (c.k > j.maxb + j.nbk) {
      c.fail |= kFailInternal;
      break;
    }
    const uint32_t ring = c.k % j.nbk;
    const uint32_t cnt = uniform(lds_ld(&w.ring[ring]));
    if (cnt == 0) {
      ++c.k;
      continue;
    }
    if (cnt > a.bcap) c.fail |= kFailBucket;
    if (c.nb >= j.maxb) c.fail |= kFailInternal;
    if (c.fail) break;
    if (lane == 0) {
      lds_st(&w.ring[ring], 0u);
      st32(&w.bnd[c.nb], c.nslot);
      st32(&w.bk[c.nb], c.k);
    }
    ++c.nb;
    uint32_t pushes = 0, lfail = 0;
    for (uint32_t base = 0; base < cnt; base += kWave) {
      const uint32_t i = base + lane;
      uint32_t v = 0, csb = 0, el = 0;
      uint64_t kv = 0;
      uint4 rw = make_uint4(0u, 0u, 0u, 0u);
      if (i < cnt) {
        const uint64_t ent = ld64(reinterpret_cast<const uint64_t *>(&w.bkt[ring * a.bcap + i]));
        v = (uint32_t)ent;
        csb = (uint32_t)(ent >> 32);
        kv = ld64(nkey(w, v));
        rw = j.rows[v];
        if constexpr (MODE == kBrandes) el = j.eligible[v];
      }
      // an entry is live while the key is still its offer: same seconds, not settled
      const bool valid = (i < cnt) && (kv & ~kLowMask) == offer_key_bits(c.tag, csb, 0u);
      const uint64_t m = __ballot(valid);
      const uint32_t cntv = (uint32_t)__popcll(m);
      if (c.nslot + cntv > a.cap) {
        c.fail |= kFailCap;
        break;
      }
      const uint32_t s = c.nslot + rank_below(m);
      c.nslot += cntv;
      uint32_t cnt_e = 0;
      if (valid) {
        vlo = v < vlo ? v : vlo;
        vhi = v > vhi ? v : vhi;
        const float secs_v = __uint_as_float(csb);
        const bool is_src = (s == 0);
        const uint32_t f0 = rw.x, nf = rw.y - rw.x, nr = rw.w - rw.z;
        st_edges += nr;
        // fold state (multi-predecessor: near candidates; trees: lexicographic min)
        const double thr_far = kMulti ? (double)secs_v * far_factor(nf) : 0.0;
        uint32_t n_near = 0, near_j = 0, near_slot = 0;
        float near_cs = INFINITY;
        float bcs = INFINITY, bsu = INFINITY;
        uint32_t bu = 0xffffffffu, bj = 0xffffffffu, bslot = 0;
        // ---- fold over the out-row, one node per lane: the candidates' loads
        // go out kFoldChunk at a time, then the candidates are weighed in row order
        constexpr uint32_t kFoldChunk = fold_chunk<MODE>();
        const bool hinted = hint && !key_marked(kv);
        const bool scan = !is_src && ((fold && !hinted) || out_canon);
        const uint32_t nfold = scan ? nf : 0u;
        for (uint32_t cb = 0; cb < nfold; cb += kFoldChunk) {
          uint2 fe[kFoldChunk];
          uint64_t fk[kFoldChunk];
#pragma unroll
          for (uint32_t q = 0; q < kFoldChunk; ++q) fe[q] = (cb + q < nf) ? j.fwd_adj[f0 + cb + q] : make_uint2(v, 0u);
#pragma unroll
          for (uint32_t q = 0; q < kFoldChunk; ++q) {
            const uint32_t u = fe[q].x & kNodeMask;
            fk[q] = 0;
            if (u != v && (fold || (cycles && (fe[q].x & kCanonBit)))) fk[q] = ld64(nkey(w, u));
          }
          // fold candidates v -> u (u settled in an earlier bucket)
#pragma unroll
          for (uint32_t q = 0; q < kFoldChunk; ++q) {
            const uint32_t u = fe[q].x & kNodeMask;
            if (u == v) continue;
            const uint64_t ku = fk[q];
            if (!key_current(ku, c.tag)) continue;
            const float su = key_secs(ku);
            if (cycles && (fe[q].x & kCanonBit) && key_less(j.orig, su, u, secs_v, v)) ++cnt_e;
            if (!fold || bucket_of(su, j.inv_delta) >= c.k) continue;
            const float cs = su + __uint_as_float(fe[q].y);
            if (cs > j.maxsec) continue;
            const uint32_t jr = cb + q;
            if constexpr (kMulti) {
              if ((double)cs <= thr_far) {
                if (n_near == 0) {
                  near_j = jr;
                  near_cs = cs;
                  near_slot = key_slot(ku);
                }
                ++n_near;
              }
            } else {
              // dijkstra_tree_shortest / dijkstra_tree_segment (centrality.rs:1186-1195,
              // 1589-1607): strict < in settle order keeps the lexicographic min
              // (cs, secs_u, u, row position)
              if (cs < bcs ||
                  (cs == bcs && (su < bsu || (su == bsu && (j.orig[u] < j.orig[bu] || (u == bu && jr < bj)))))) {
                bcs = cs;
                bsu = su;
                bu = u;
                bj = jr;
                bslot = key_slot(ku);
              }
            }
          }
        }
        // settle record
        if constexpr (kMulti) {
          const bool wide = nf > kWideRow;
          Fold f;
          f.sigma = 1.0;
          f.pmask = 0;
          f.wide = wide;
          f.overflow = false;
          uint32_t pslot0 = 0, npred = 0;
          if (!is_src && fold) {
            if (hinted) {
              // no tie mark: the one offer within the band is the node's final
              // seconds, and its key names the offering slot
              pslot0 = key_slot(kv);
              f.sigma = ldd(&w.sig[pslot0]);
              npred = 1;
            } else if (n_near == 1) {
              f.pmask = ps_single(near_j, wide);
              f.sigma = ldd(&w.sig[near_slot]);
              pslot0 = near_slot;
              npred = 1;
              if (near_cs != secs_v) lfail |= kFailInternal;
            } else if (n_near == 0 ||
                       !brandes_fold_exact(j, w, c.tag, v, secs_v, c.k, f0, f0 + nf, thr_far, n_near, f)) {
              lfail |= f.overflow ? kFailWide : kFailInternal;
            } else {
              npred = ps_count(f.pmask, wide);
              const uint32_t u0 = j.fwd_adj[f0 + ps_first(f.pmask, wide)].x & kNodeMask;
              pslot0 = key_slot(ld64(nkey(w, u0)));
            }
          }
          // the node's key turns settled with its own slot and the slot takes its
          // sigma (read by folds, Phase 2 and back-propagation, none of which
          // runs for closeness alone); a settled key outranks every offer
          if (MODE != kBrandes || (j.flags & 2u)) {
            st64(nkey(w, v), settle_key(kv, s));
            std_(&w.sig[s], f.sigma);
          }
          // the slot record is read only after this bucket's fence: one plain 16-byte store
          *reinterpret_cast<uint4 *>(&w.rec[2 * (size_t)s]) =
              make_uint4(v, csb, pslot0, npred | (el ? 0x100u : 0u));
          // the predecessor mask is needed only to walk several predecessors
          if (MODE == kBrandesExport || npred > 1) st64(&w.pm[s], f.pmask);
          if constexpr (MODE == kBrandes) {
            const float route_v = secs_v * speed;
            if (j.flags & 8u) std_(&w.ods[s], 0.0);  // set below for the source's destinations
            if (j.flags & 2u) {
              // dependencies for thresholds the node lies beyond are never
              // accumulated (successors lie farther still) nor read
#pragma unroll
              for (int t = 0; t < NT; ++t) {
                if (route_v <= j.dist[t]) {
                  std_(dep_at(w, t, s), 0.0);
                  std_(dep_at(w, NT + t, s), 0.0);
                }
              }
            }
          }
        } else {
          uint32_t pslot = 0xffffffffu;
          float olen = 0.0f, llen = 0.0f;
          if (!is_src && MODE != kAccess) {
            if (bj == 0xffffffffu || bcs != secs_v) {
              lfail |= kFailInternal;
              bj = 0xffffffffu;
            } else {
              pslot = bslot;
              llen = j.fwd_twlen[f0 + bj];
              olen = (pslot == 0) ? llen : slot_olen(w, pslot);
            }
          }
          st64(nkey(w, v), settle_key(kv, s));
          st64(&w.rec[2 * (size_t)s], ((uint64_t)csb << 32) | v);
          if constexpr (MODE != kAccess) {
            st64(&w.rec[2 * (size_t)s + 1], ((uint64_t)bj << 32) | pslot);
            st64(&w.pm[s], ((uint64_t)__float_as_uint(olen) << 32) | __float_as_uint(llen));
          }
          if constexpr (MODE == kSegment) {
#pragma unroll
            for (int t = 0; t < NT; ++t) std_(&w.acc[(size_t)s * NT + t], 0.0);
          }
        }
      }
      // ---- relaxations of this batch's settled nodes, one in-edge per lane: an
      // exclusive scan of the in-degrees gives each node's first edge; the
      // owner of edge g is the last lane whose first edge is <= g. The old key
      // returned by the atomic max also tells whether the neighbour settled
      // before the node (canonical edges count toward the circuit rank).
      // A plain load first: keys only grow within an epoch, so a neighbour whose
      // key already beats the offer (settled nodes, and better offers) needs no
      // atomic, which would leave a dirty line behind for nothing. The decisions
      // taken from the loaded key are the atomic's: a neighbour settling before
      // the node is final when its bucket opens. Every offer meets the best
      // offer before it (loaded or returned by the atomic); when the loser lies
      // within the widest tie band of the winner, the neighbour is tie-marked
      // and folds over its whole out-row. Two offers within the band of the
      // final seconds always meet this way, whichever arrives first.
      {
        const uint32_t nrl = valid ? rw.w - rw.z : 0u;
        const uint32_t first = __ockl_wfscan_add_u32(nrl, false);
        const uint32_t tot = uniform(wave_sum(nrl));
        if (tot) {
          lds_st(&w.rel[lane], first);
          lds_st(&w.rel[kWave + lane], nrl);
          lds_st(&w.rel[2 * kWave + lane], v);
          lds_st(&w.rel[3 * kWave + lane], csb);
          lds_st(&w.rel[4 * kWave + lane], rw.z);
          lds_st(&w.rel[5 * kWave + lane], s);
          wave_fence();
          for (uint32_t g0 = 0; g0 < tot; g0 += kRelQ * kWave) {
            // kRelQ edges per lane per step: owners and adjacency first, then the
            // neighbours' keys, then the atomics, each stage's loads in flight together
            bool act[kRelQ], ce[kRelQ];
            uint32_t vo[kRelQ], oslot[kRelQ];
            float so[kRelQ], route_o[kRelQ];
            uint2 re[kRelQ];
#pragma unroll
            for (uint32_t q = 0; q < kRelQ; ++q) {
              const uint32_t g = g0 + q * kWave + lane;
              act[q] = false;
              ce[q] = false;
              route_o[q] = 0.0f;
              vo[q] = oslot[q] = 0;
              so[q] = 0.0f;
              re[q] = make_uint2(0u, 0u);
              if (g < tot) {
                uint32_t lo = 0, hi = kWave - 1;
                while (lo < hi) {
                  const uint32_t mid = (lo + hi + 1) >> 1;
                  if (lds_ld(&w.rel[mid]) <= g) lo = mid;
                  else hi = mid - 1;
                }
                const uint32_t k = g - lds_ld(&w.rel[lo]);
                if (k >= lds_ld(&w.rel[kWave + lo])) {
                  lfail |= kFailInternal;
                } else {
                  vo[q] = lds_ld(&w.rel[2 * kWave + lo]);
                  so[q] = __uint_as_float(lds_ld(&w.rel[3 * kWave + lo]));
                  oslot[q] = lds_ld(&w.rel[5 * kWave + lo]);
                  re[q] = j.rev_adj[lds_ld(&w.rel[4 * kWave + lo]) + k];
                  act[q] = true;
                }
              }
            }
            uint64_t k0[kRelQ];
            float csq[kRelQ];
            bool offer[kRelQ];
#pragma unroll
            for (uint32_t q = 0; q < kRelQ; ++q) {
              const uint32_t nb = re[q].x & kNodeMask;
              act[q] = act[q] && nb != vo[q];
              csq[q] = so[q] + __uint_as_float(re[q].y);
              offer[q] = act[q] && csq[q] <= j.maxsec;
              k0[q] = (offer[q] || (act[q] && cycles && (re[q].x & kCanonBit))) ? ld64(nkey(w, nb)) : 0ull;
            }
#pragma unroll
            for (uint32_t q = 0; q < kRelQ; ++q) {
              if (!act[q]) continue;
              const uint32_t nb = re[q].x & kNodeMask;
              const float cs = csq[q];
              uint64_t kn = k0[q];
              if (offer[q]) {
                const uint64_t nk = offer_key(c.tag, cs, oslot[q]);
                kn = (k0[q] >= nk) ? k0[q] : wmax64(nkey(w, nb), nk);
                const bool better = (nk >> 23) > (kn >> 23);  // strictly fewer seconds (or an older epoch)
                if (hint && key_current(kn, c.tag) && !key_settled(kn)) {
                  const double sk = (double)key_secs(kn);
                  const double lose = better ? sk : (double)cs, win = better ? (double)cs : sk;
                  if (lose <= win * j.far_max)
                    __hip_atomic_fetch_or(nkey(w, nb), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                }
                if (better) {
                  const uint32_t b = bucket_of(cs, j.inv_delta);
                  if (b <= c.k || b - c.k >= j.nbk) {
                    lfail |= (b <= c.k) ? kFailInternal : kFailRing;
                  } else {
                    const uint32_t rb = b % j.nbk;
                    const uint32_t pos = lds_add(&w.ring[rb], 1u);
                    if (pos < a.bcap)
                      st64(reinterpret_cast<uint64_t *>(&w.bkt[rb * a.bcap + pos]),
                           ((uint64_t)__float_as_uint(cs) << 32) | nb);
                    else
                      lfail |= kFailBucket;
                    ++pushes;
                  }
                }
              }
              if (cycles && (re[q].x & kCanonBit) && key_current(kn, c.tag) &&
                  key_less(j.orig, key_secs(kn), nb, so[q], vo[q])) {
                ce[q] = true;
                route_o[q] = so[q] * speed;
              }
            }
            if constexpr (MODE == kBrandes) {
              if (cycles) {
#pragma unroll
                for (uint32_t q = 0; q < kRelQ; ++q)
#pragma unroll
                  for (int t = 0; t < NT; ++t) eband[t] += (uint32_t)__popcll(__ballot(ce[q] && route_o[q] <= j.dist[t]));
              }
            }
          }
          wave_fence();  // the next batch rewrites the LDS table
        }
      }
      // nodes within each threshold (circuit rank, reach; segment: the count
      // of node-threshold pairs the statistics report) and the rank's edge
      // counts from the out-row side, reduced over the wave
      if constexpr (MODE == kBrandes || MODE == kSegment) {
        const float route_v = __uint_as_float(csb) * speed;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const bool in = valid && route_v <= j.dist[t];
          nband[t] += (uint32_t)__popcll(__ballot(in));
          if (cycles) eband[t] += uniform(wave_sum(in ? cnt_e : 0u));
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
    for (uint32_t r = lane; r < j.nbk; r += kWave) lds_st(&w.ring[r], 0u);
    wave_fence();
    return false;
  }
  if (lane == 0) st32(&w.bnd[c.nb], c.nslot);
  c.vlo = uniform(__ockl_wfred_min_u32(vlo));
  c.vhi = uniform(__ockl_wfred_max_u32(vhi));
  wave_fence();
  return true;
}

// Next item for a wave on XCD `x`: its own share first, then the other XCDs'
// shares in turn. Returns n_items when all are taken.
__device__ __forceinline__ uint32_t next_item(const SearchArgs &a, uint32_t x) {
  uint32_t si = a.n_items;
  if (lane_id() == 0) {
    for (uint32_t q = 0; q < kXcd; ++q) {
      const uint32_t xx = (x + q) % kXcd;
      const uint32_t len = a.xbeg[xx + 1] - a.xbeg[xx];
      if (__hip_atomic_load(&a.next[xx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= len) continue;
      const uint32_t i = atomicAdd(&a.next[xx], 1u);
      if (i < len) {
        si = a.xbeg[xx] + i;
        break;
      }
    }
  }
  return uniform(si);
}

// Phase 2 of the multi-predecessor fold (centrality.rs:1443-1482): with a
// tolerance above the internal tie band, predecessors and sigma are rebuilt
// from the final distances. u is a predecessor of v when u settles before v
// and secs_u + es(v->u) <= secs_v * (1 + tol) (one entry per distinct u);
// sigma_v sums sigma_u in u's settle order. A predecessor can settle in v's
// own bucket, so each bucket resolves in rounds: a node is ready once all its
// predecessors are. Record bit 0x200 of the metadata word marks a resolved
// node; bits 16-31 hold the round in which it resolved, its level: every
// same-bucket predecessor has a lower level, so back-propagation can visit a
// bucket level by level from the top. Returns false if a bucket fails to
// resolve (cannot happen: every predecessor settles strictly earlier).
constexpr uint32_t kP2Done = 0x200u;

__device__ bool phase2_rebuild(CJob &j, const WaveWs &w, const SrcCtx &c, uint32_t &fail) {
  const uint32_t lane = lane_id();
  const float band = 1.0f + j.tol;
  bool overflow = false;
  for (uint32_t b = 0; b < c.nb; ++b) {
    const uint32_t s0 = uniform(ld32(&w.bnd[b])), s1 = uniform(ld32(&w.bnd[b + 1]));
    // predecessor masks over the out-row positions of each node
    for (uint32_t s = s0 + lane; s < s1; s += kWave) {
      const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
      const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
      const uint32_t v = (uint32_t)lo;
      const float secs_v = __uint_as_float((uint32_t)(lo >> 32));
      uint32_t meta = (uint32_t)(hi >> 32) & 0x100u;  // keep the eligibility bit
      uint64_t pm = 0;
      if (s == 0) {
        meta |= kP2Done;
        std_(&w.sig[0], 1.0);
      } else {
        const uint4 rw = j.rows[v];
        const bool wide = (rw.y - rw.x) > kWideRow;
        const float lim = secs_v * band;
        for (uint32_t p = rw.x; p < rw.y; ++p) {
          const uint2 e = j.fwd_adj[p];
          const uint32_t u = e.x & kNodeMask;
          if (u == v) continue;
          const uint64_t ku = ld64(nkey(w, u));
          if (!key_current(ku, c.tag)) continue;
          const float su = key_secs(ku);
          if (!key_less(j.orig, su, u, secs_v, v)) continue;
          const float ps = su + __uint_as_float(e.y);
          if (!(ps <= lim)) continue;
          bool dup = false;
          uint32_t q;
          for (uint64_t mm = pm; ps_next(mm, wide, q);)
            if ((j.fwd_adj[rw.x + q].x & kNodeMask) == u) dup = true;
          if (!dup) pm = ps_add(pm, p - rw.x, wide, overflow);
        }
      }
      st64(&w.pm[s], pm);
      st64(&w.rec[2 * (size_t)s + 1], ((uint64_t)meta << 32) | (uint32_t)hi);
    }
    if (uniform(wave_sum((uint32_t)overflow))) {
      fail |= kFailWide;
      return false;
    }
    wave_fence();
    for (uint32_t round = 0;; ++round) {
      uint32_t pending = 0;
      for (uint32_t s = s0 + lane; s < s1; s += kWave) {
        const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
        const uint32_t meta = (uint32_t)(hi >> 32);
        if (meta & kP2Done) continue;
        const uint32_t v = (uint32_t)ld64(&w.rec[2 * (size_t)s]);
        const uint64_t pm = ld64(&w.pm[s]);
        const uint4 rv = j.rows[v];
        const uint32_t f0 = rv.x;
        const bool wide = (rv.y - rv.x) > kWideRow;
        bool ready = true;
        uint32_t q;
        for (uint64_t mm = pm; ps_next(mm, wide, q);) {
          const uint32_t u = j.fwd_adj[f0 + q].x & kNodeMask;
          const uint32_t su = key_slot(ld64(nkey(w, u)));
          const uint32_t mu = (uint32_t)(ld64(&w.rec[2 * (size_t)su + 1]) >> 32);
          // same-bucket predecessors must have resolved in an earlier round
          if (!(mu & kP2Done) || (su >= s0 && su < s1 && (mu >> 16) >= round)) ready = false;
        }
        if (!ready) {
          ++pending;
          continue;
        }
        // sigma over the predecessors in their settle order (secs, node index)
        double sigma = 0.0;
        uint32_t pslot0 = 0;
        bool have_last = false;
        float lsu = 0.0f;
        uint32_t lu = 0;
        const uint32_t np = ps_count(pm, wide);
        for (uint32_t step = 0; step < np; ++step) {
          bool found = false;
          float bsu = 0.0f;
          uint32_t bu = 0, bsl = 0;
          for (uint64_t mm = pm; ps_next(mm, wide, q);) {
            const uint32_t u = j.fwd_adj[f0 + q].x & kNodeMask;
            const uint64_t ku = ld64(nkey(w, u));
            const float su = key_secs(ku);
            if (have_last && !key_less(j.orig, lsu, lu, su, u)) continue;
            if (!found || key_less(j.orig, su, u, bsu, bu)) {
              found = true;
              bsu = su;
              bu = u;
              bsl = key_slot(ku);
            }
          }
          sigma += ldd(&w.sig[bsl]);
          have_last = true;
          lsu = bsu;
          lu = bu;
        }
        if (pm) pslot0 = key_slot(ld64(nkey(w, (j.fwd_adj[f0 + ps_first(pm, wide)].x & kNodeMask))));
        std_(&w.sig[s], sigma);
        st64(&w.rec[2 * (size_t)s + 1], ((uint64_t)(meta | kP2Done | np | (round << 16)) << 32) | pslot0);
      }
      pending = uniform(wave_sum(pending));
      if (pending == 0 && lane_id() == 0) st32(&w.lvl[b], round);
      wave_fence();
      if (pending == 0) break;
      if (round > s1 - s0 || round >= 0xffffu) return false;
    }
  }
  return true;
}

// Phase 1 folded after the search (flag 16), for bucket widths that cannot
// keep a node's tie band clear of its own bucket (thresholds long against the
// shortest edge). The search settles every node at its exact seconds; here,
// bucket by bucket, each node's predecessors and sigma come from the ordered
// fold over all its out-row candidates settled before it (in seconds, original
// index order), same-bucket ones included, once every such candidate has been
// folded itself: in rounds, as Phase 2 does, whose numbers become the levels
// back-propagation walks.
__device__ bool phase1_fold_after(CJob &j, const WaveWs &w, const SrcCtx &c, uint32_t &fail) {
  const uint32_t lane = lane_id();
  bool overflow = false;
  for (uint32_t b = 0; b < c.nb; ++b) {
    const uint32_t s0 = uniform(ld32(&w.bnd[b])), s1 = uniform(ld32(&w.bnd[b + 1]));
    // candidate masks over the out-row positions of each node
    for (uint32_t s = s0 + lane; s < s1; s += kWave) {
      const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
      const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
      const uint32_t v = (uint32_t)lo;
      const float secs_v = __uint_as_float((uint32_t)(lo >> 32));
      uint32_t meta = (uint32_t)(hi >> 32) & 0x100u;  // keep the eligibility bit
      uint64_t cm = 0;
      if (s == 0) {
        meta |= kP2Done;
        std_(&w.sig[0], 1.0);
      } else {
        const uint4 rw = j.rows[v];
        const bool wide = (rw.y - rw.x) > kWideRow;
        const double thr_far = (double)secs_v * far_factor(rw.y - rw.x);
        for (uint32_t p = rw.x; p < rw.y; ++p) {
          const uint2 e = j.fwd_adj[p];
          const uint32_t u = e.x & kNodeMask;
          if (u == v) continue;
          const uint64_t ku = ld64(nkey(w, u));
          if (!key_current(ku, c.tag)) continue;
          const float su = key_secs(ku);
          if (!key_less(j.orig, su, u, secs_v, v)) continue;
          const float cs = su + __uint_as_float(e.y);
          if (cs > j.maxsec || (double)cs > thr_far) continue;
          cm = ps_add(cm, p - rw.x, wide, overflow);
        }
      }
      st64(&w.pm[s], cm);
      st64(&w.rec[2 * (size_t)s + 1], ((uint64_t)meta << 32) | (uint32_t)hi);
    }
    if (uniform(wave_sum((uint32_t)overflow))) {
      fail |= kFailWide;
      return false;
    }
    wave_fence();
    for (uint32_t round = 0;; ++round) {
      uint32_t pending = 0, bad = 0;
      for (uint32_t s = s0 + lane; s < s1; s += kWave) {
        const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
        const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
        const uint32_t meta = (uint32_t)(hi >> 32);
        if (meta & kP2Done) continue;
        const uint32_t v = (uint32_t)lo;
        const float secs_v = __uint_as_float((uint32_t)(lo >> 32));
        const uint64_t cm = ld64(&w.pm[s]);
        const uint4 rw = j.rows[v];
        const bool wide = (rw.y - rw.x) > kWideRow;
        bool ready = true;
        uint32_t q;
        for (uint64_t mm = cm; ps_next(mm, wide, q);) {
          const uint32_t u = j.fwd_adj[rw.x + q].x & kNodeMask;
          const uint32_t su = key_slot(ld64(nkey(w, u)));
          const uint32_t mu = (uint32_t)(ld64(&w.rec[2 * (size_t)su + 1]) >> 32);
          // same-bucket candidates must have been folded in an earlier round
          if (!(mu & kP2Done) || (su >= s0 && su < s1 && (mu >> 16) >= round)) ready = false;
        }
        if (!ready) {
          ++pending;
          continue;
        }
        Fold f;
        f.overflow = false;
        const uint32_t n_near = ps_count(cm, wide);
        const double thr_far = (double)secs_v * far_factor(rw.y - rw.x);
        if (n_near == 0 || !brandes_fold_exact(j, w, c.tag, v, secs_v, 0u, rw.x, rw.y, thr_far, n_near, f, true)) {
          if (f.overflow) overflow = true;
          ++bad;
          continue;
        }
        const uint32_t np = ps_count(f.pmask, wide);
        const uint32_t u0 = j.fwd_adj[rw.x + ps_first(f.pmask, wide)].x & kNodeMask;
        const uint32_t pslot0 = key_slot(ld64(nkey(w, u0)));
        st64(&w.pm[s], f.pmask);
        std_(&w.sig[s], f.sigma);
        st64(&w.rec[2 * (size_t)s + 1], ((uint64_t)(meta | kP2Done | np | (round << 16)) << 32) | pslot0);
      }
      if (uniform(wave_sum(bad))) {
        if (uniform(wave_sum((uint32_t)overflow))) fail |= kFailWide;
        return false;
      }
      pending = uniform(wave_sum(pending));
      if (pending == 0 && lane_id() == 0) st32(&w.lvl[b], round);
      wave_fence();
      if (pending == 0) break;
      if (round > s1 - s0 || round >= 0xffffu) return false;
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// Results of one source (centrality.rs:1733-1780 closeness, 1783-1835
// betweenness credit): every reached target but the source adds, per threshold
// it lies within, its five closeness terms and its two betweenness credits to
// the shared accumulators.
//
// The settled targets are visited in node-index order where that is cheap: the
// wave scans the epoch-tagged keys over [vlo, vhi], compacts the reached ones
// (node, seconds, slot) into an LDS list and hands lane i the i-th, so the
// lanes of one atomic instruction hit neighbouring addresses of a metric row
// (a few cache lines) rather than 64 scattered ones. The route comes from the
// key and the eligibility from the graph, so closeness reads no slot record.
// A sparse span falls back to slot order.
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void flush_target(CJob &j, const WaveWs &w, uint32_t v, uint32_t secs_bits, uint32_t s,
                                             const float (&rank)[NT], float wt, float cyc) {
  const float route = __uint_as_float(secs_bits) * j.speed;
  const size_t ts = j.tstride, nacc = j.nacc;
  double *accv = j.acc + v;
  // the dependencies are loaded before any atomic is issued: vmcnt counts
  // stores and atomics too, in order, so a load issued behind them would
  // wait for every one of them to reach L2
  double dpl[NT], dbt[NT];
  double pc = 0.0;
  if (j.flags & 2u) {
    pc = (j.flags & 8u) ? ldd(&w.ods[s]) : (j.eligible[v] ? 0.5 : 1.0);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      dpl[t] = dbt[t] = 0.0;
      if (route <= j.dist[t]) {
        dpl[t] = ldd(dep_at(w, t, s));
        dbt[t] = ldd(dep_at(w, NT + t, s));
      }
    }
  }
  if (j.flags & 1u) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (route <= j.dist[t]) {
        gadd_f64(&accv[((size_t)0 * ts + t) * nacc], (double)wt);
        gadd_f64(&accv[((size_t)1 * ts + t) * nacc], (double)(route * wt));
        gadd_f64(&accv[((size_t)2 * ts + t) * nacc], (double)(rank[t] * cyc));
        gadd_f64(&accv[((size_t)3 * ts + t) * nacc], (double)((1.0f / route) * wt));
        gadd_f64(&accv[((size_t)4 * ts + t) * nacc], (double)(exp_f32(-j.beta[t] * route) * wt));
      }
    }
  }
  if (j.flags & 2u) {
    // seed pc: the all-pairs weight (0.5 between two eligible sources), or the OD trip weight
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (route <= j.dist[t]) {
        // dependency = seed + delta; credit = dependency - seed (centrality.rs:848-866)
        const double seed_b = pc * exp(-(double)j.beta[t] * (double)route);
        const double credit = (pc + dpl[t]) - pc;
        const double credit_b = (seed_b + dbt[t]) - seed_b;
        if (credit > 0.0) gadd_f64(&accv[((size_t)5 * ts + t) * nacc], credit * (double)wt);
        if (credit_b > 0.0) gadd_f64(&accv[((size_t)6 * ts + t) * nacc], credit_b * (double)wt);
      }
    }
  }
}

template <int NT>
__device__ void flush_targets(CJob &j, const WaveWs &w, const SrcCtx &c, uint32_t R, const float (&rank)[NT],
                              float wt, float cyc) {
  if (R <= 1) return;
  const uint32_t lane = lane_id();
  const uint32_t lo = c.vlo, hi = c.vhi;
  if ((uint64_t)(hi - lo) + 1 > (uint64_t)kScanSpan * R) {
    for (uint32_t s = 1 + lane; s < R; s += kWave) {
      const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
      flush_target<NT>(j, w, (uint32_t)lo, (uint32_t)(lo >> 32), s, rank, wt, cyc);
    }
    return;
  }
  for (uint64_t base = lo; base <= hi; base += (uint64_t)kScanQ * kWave) {
    uint64_t key[kScanQ];
#pragma unroll
    for (uint32_t q = 0; q < kScanQ; ++q) {
      const uint64_t v = base + q * kWave + lane;
      key[q] = (v <= hi) ? ld64(nkey(w, v)) : 0ull;
    }
    uint32_t n = 0;
#pragma unroll
    for (uint32_t q = 0; q < kScanQ; ++q) {
      const uint32_t v = (uint32_t)(base + q * kWave + lane);
      const bool hit = key_current(key[q], c.tag) && v != c.src;
      const uint64_t m = __ballot(hit);
      if (hit) {
        const uint32_t i = n + rank_below(m);
        lds_st(&w.list[i], v);
        lds_st(&w.list[kFlushList + i], __float_as_uint(key_secs(key[q])));
        lds_st(&w.list[2 * kFlushList + i], key_slot(key[q]));
      }
      n += (uint32_t)__popcll(m);
    }
    wave_fence();
    for (uint32_t i = lane; i < n; i += kWave)
      flush_target<NT>(j, w, lds_ld(&w.list[i]), lds_ld(&w.list[kFlushList + i]), lds_ld(&w.list[2 * kFlushList + i]),
                       rank, wt, cyc);
    wave_fence();
  }
}

// ---------------------------------------------------------------------------
// Persistent kernel: each wave loops over sources taken from the counters.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Land-use statistics of one source (data.rs:946-1233), computed by its wave
// from the list of entries it reaches (one per dedupe group, with its total
// distance), in ascending entry order: f32 sums in that order, medians over
// the values sorted by Rust's f32 total order (ties kept in list order, as
// the reference's stable sort keeps them).

// f32 total order (f32::total_cmp) as an ascending u32 key.
__device__ __forceinline__ uint32_t total_key(float x) {
  const uint32_t b = __float_as_uint(x);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

// Sorts buf[0..n) ascending in the wave's scratch (room for the next power of two).
__device__ void wave_sort64(uint64_t *buf, uint32_t n) {
  const uint32_t lane = lane_id();
  uint32_t P = 1;
  while (P < n) P <<= 1;
  for (uint32_t i = n + lane; i < P; i += kWave) st64(&buf[i], ~0ull);
  wave_fence();
  for (uint32_t k = 2; k <= P; k <<= 1) {
    for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
      for (uint32_t i = lane; i < P; i += kWave) {
        const uint32_t l = i ^ jj;
        if (l > i) {
          const uint64_t x = ld64(&buf[i]), y = ld64(&buf[l]);
          if ((x > y) == ((i & k) == 0)) {
            st64(&buf[i], y);
            st64(&buf[l], x);
          }
        }
      }
      wave_fence();
    }
  }
}

// weighted_median (data.rs:961-997) over keys sorted by value, the value and
// weight of key k being val(k) and wts[k]; one lane.
template <typename V>
__device__ float weighted_median_sorted(const uint64_t *keys, uint32_t n, const float *wts, float total_wt, V val) {
  if (n == 0 || total_wt == 0.0f) return __int_as_float(0x7fc00000);
  const float mid = total_wt / 2.0f;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = (uint32_t)ld64(&keys[i]);
    if (wts[k] > mid) return val(k);
  }
  float agg = 0.0f;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t k = (uint32_t)ld64(&keys[i]);
    agg = agg + wts[k];
    if (agg == mid) return i + 1 < n ? (val(k) + val((uint32_t)ld64(&keys[i + 1]))) / 2.0f : val(k);
    if (agg > mid) return val(k);
  }
  return val((uint32_t)ld64(&keys[n - 1]));
}

// median (data.rs:946-959) of keys sorted by value.
template <typename V>
__device__ __forceinline__ float median_sorted(const uint64_t *keys, uint32_t n, V val) {
  if (n == 0) return __int_as_float(0x7fc00000);
  if (n & 1u) return val((uint32_t)ld64(&keys[n / 2]));
  return (val((uint32_t)ld64(&keys[n / 2 - 1])) + val((uint32_t)ld64(&keys[n / 2]))) / 2.0f;
}

// The fourteen statistics of every map and threshold for one source, from
// its winners win[0..nw) = {entry << 32 | bits(total distance)}.
__device__ void source_stats(const AccessOut &ao, uint32_t src, uint64_t *win, uint32_t nw, uint64_t *keys,
                             float *vals, float *wts) {
  const uint32_t lane = lane_id();
  const float nan = __int_as_float(0x7fc00000);
  wave_sort64(win, nw);  // ascending entry
  for (uint32_t mp = 0; mp < ao.n_maps; ++mp) {
    const float *num_of = ao.values + (size_t)mp * ao.n_entries;
    for (uint32_t t = 0; t < ao.T; ++t) {
      // the values within this threshold, in entry order
      uint32_t n = 0;
      for (uint32_t base = 0; base < nw; base += kWave) {
        const uint32_t i = base + lane;
        bool keep = false;
        float num = 0.0f, wt = 0.0f;
        if (i < nw) {
          const uint64_t wv = ld64(&win[i]);
          const float tot = __uint_as_float((uint32_t)wv);
          num = num_of[(uint32_t)(wv >> 32)];
          keep = tot <= ao.dist[t] && !isnan(num);
          if (keep) wt = __fdiv_rn(fminf(exp_f32(__fmul_rn(-ao.beta[t], tot)), ao.mcw[t]), ao.mcw[t]);
        }
        const uint64_t bal = __ballot(keep);
        if (keep) {
          const uint32_t k = n + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
          vals[k] = num;
          wts[k] = wt;
        }
        n += (uint32_t)__popcll(bal);
      }
      wave_fence();
      // sums, counts, extremes (one lane, list order)
      float r[kStatMetrics];
      float count_wt = 0.0f, median = nan, median_wt = nan;
      if (lane == 0) {
        float sum = 0.0f, sum_wt = 0.0f, count = 0.0f, sum_sq = 0.0f, sum_sq_wt = 0.0f, mn = nan, mx = nan;
        for (uint32_t k = 0; k < n; ++k) {
          const float v = vals[k], w = wts[k];
          sum = sum + v;
          sum_wt = sum_wt + v * w;
          count = count + 1.0f;
          count_wt = count_wt + w;
          sum_sq = sum_sq + v * v;
          sum_sq_wt = sum_sq_wt + w * v * v;
          mx = isnan(mx) ? v : fmaxf(mx, v);
          mn = isnan(mn) ? v : fminf(mn, v);
        }
        const float mean = count > 0.0f ? sum / count : nan;
        const float mean_wt = count_wt > 0.0f ? sum_wt / count_wt : nan;
        r[0] = sum;
        r[1] = sum_wt;
        r[2] = mean;
        r[3] = mean_wt;
        r[6] = count;
        r[7] = count_wt;
        r[8] = count > 0.0f ? fmaxf(sum_sq / count - mean * mean, 0.0f) : nan;
        r[9] = count_wt > 0.0f ? fmaxf(sum_sq_wt / count_wt - mean_wt * mean_wt, 0.0f) : nan;
        r[12] = mx;
        r[13] = mn;
      }
      count_wt = __shfl(count_wt, 0);
      auto value = [&](uint32_t k) { return vals[k]; };
      // median and weighted median: the values in total order
      for (uint32_t k = lane; k < n; k += kWave) st64(&keys[k], ((uint64_t)total_key(vals[k]) << 32) | k);
      wave_fence();
      wave_sort64(keys, n);
      if (lane == 0) {
        median = median_sorted(keys, n, value);
        median_wt = weighted_median_sorted(keys, n, wts, count_wt, value);
        r[4] = median;
        r[5] = median_wt;
      }
      median = __shfl(median, 0);
      median_wt = __shfl(median_wt, 0);
      wave_fence();
      // median absolute deviation
      auto dev = [&](uint32_t k) { return fabsf(vals[k] - median); };
      for (uint32_t k = lane; k < n; k += kWave) st64(&keys[k], ((uint64_t)total_key(dev(k)) << 32) | k);
      wave_fence();
      wave_sort64(keys, n);
      if (lane == 0) r[10] = (n > 0 && !isnan(median)) ? median_sorted(keys, n, dev) : nan;
      wave_fence();
      // weighted median absolute deviation
      auto dev_wt = [&](uint32_t k) { return fabsf(vals[k] - median_wt); };
      for (uint32_t k = lane; k < n; k += kWave) st64(&keys[k], ((uint64_t)total_key(dev_wt(k)) << 32) | k);
      wave_fence();
      wave_sort64(keys, n);
      if (lane == 0) {
        r[11] = (n > 0 && !isnan(median_wt) && count_wt > 0.0f) ? weighted_median_sorted(keys, n, wts, count_wt, dev_wt)
                                                                 : nan;
        for (uint32_t m = 0; m < kStatMetrics; ++m)
          ao.stats[(((size_t)mp * kStatMetrics + m) * ao.tstride + t) * ao.N + src] = (double)r[m];
      }
      wave_fence();
    }
  }
}

// ---- data entries reachable from a source (data.rs:520-615, 680-741), one
// wave. reach(s) gives the source's s-th reached node as {node | seconds << 32}
// (s < R). A member (entry e at node v, data distance dd) counts when v lies
// below the walk limit, secs_v * speed + dd is within the largest distance and
// no other member of its dedupe group is nearer (equal totals: the lower member
// index wins). Groups of several members keep two epoch-tagged words each in
// grec: the nearest total (pass 1), then the lowest member index at that total
// (pass 2), so each group counts once, through its nearest member, in passes
// linear in the members reached. `tag` is the source's epoch in the top bits,
// above every tag grec holds from earlier sources. Counts go to the rows of
// `ao`; with ao.stats the winners go to the list at stb (count, then up to
// st_cap winners and the scratch of source_stats) and become statistics rows.
// Returns false when a statistics source has more winners than st_cap.
template <typename Reach>
__device__ bool access_fold(const AccessOut &ao, uint32_t src, uint32_t R, const Reach &reach, float maxsec,
                            float speed, uint64_t *grec, uint64_t tag, char *stb, uint32_t st_cap) {
  const uint32_t lane = lane_id();
  uint32_t *st_n = reinterpret_cast<uint32_t *>(stb);
  uint64_t *st_win = reinterpret_cast<uint64_t *>(stb + 256);
  if (ao.stats && lane == 0) st32(st_n, 0u);
  for (int pass = 0; pass < 3; ++pass) {
    if (pass < 2 && ao.n_shared == 0) continue;
    for (uint32_t s = lane; s < R; s += kWave) {
      const uint64_t lo = reach(s);
      const uint32_t v = (uint32_t)lo;
      const float secs = __uint_as_float((uint32_t)(lo >> 32));
      if (!(secs < maxsec)) continue;
      const float net = __fmul_rn(secs, speed);
      for (uint32_t k = ao.node_off[v]; k < ao.node_off[v + 1]; ++k) {
        const uint32_t m = ao.node_mem[k];
        const uint4 me = ao.mem[m];
        const float tot = __fadd_rn(net, __uint_as_float(me.z));
        if (!(tot <= ao.maxdist)) continue;
        const bool shared = me.w != kSoloGroup;
        const uint64_t tkey = tag | ((uint64_t)(uint32_t)~__float_as_uint(tot) << 22);  // larger = nearer
        if (pass == 0) {
          if (shared) wmax64(&grec[2 * (size_t)me.w], tkey);
          continue;
        }
        if (pass == 1) {
          if (shared && ld64(&grec[2 * (size_t)me.w]) == tkey) wmax64(&grec[2 * (size_t)me.w + 1], tag | (uint32_t)~m);
          continue;
        }
        if (shared && ld64(&grec[2 * (size_t)me.w + 1]) != (tag | (uint32_t)~m)) continue;
        if (ao.stats) {
          const uint32_t q = __hip_atomic_fetch_add(st_n, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          if (q < st_cap) st64(&st_win[q], ((uint64_t)me.y << 32) | __float_as_uint(tot));
          continue;
        }
        const int32_t cls = ao.entry_class[me.y];
        if (cls < 0) continue;
        for (uint32_t t = 0; t < ao.T; ++t) {
          if (!(tot <= ao.dist[t])) continue;
          const size_t row = ((size_t)cls * ao.tstride + t) * ao.N + src;
          gadd_f64(&ao.count[row], 1.0);
          if (ao.weighted) {  // mixed uses read only the counts
            // clipped_beta_wt (common.rs:339-348): min(exp(-beta d), mcw) / mcw in f32
            const float raw = exp_f32(__fmul_rn(-ao.beta[t], tot));
            gadd_f64(&ao.weighted[row], (double)__fdiv_rn(fminf(raw, ao.mcw[t]), ao.mcw[t]));
          }
          if (ao.max_mask & (1u << t))
            atomicMin(&ao.nearest[(size_t)cls * ao.N + src], (unsigned long long)__double_as_longlong((double)tot));
        }
      }
    }
    wave_fence();
  }
  if (ao.stats) {
    const uint32_t nw = uniform(ld32(st_n));
    if (nw > st_cap) return false;
    uint64_t *st_keys = st_win + 2 * (size_t)st_cap;
    float *st_vals = reinterpret_cast<float *>(st_keys + 2 * (size_t)st_cap);
    source_stats(ao, src, st_win, nw, st_keys, st_vals, st_vals + st_cap);
  }
  return true;
}

// With up to four thresholds the search asks for five resident waves per SIMD
// (96 VGPRs, a few spilled to scratch in cold paths): 45.2 ms against 45.6 ms at
// four on the 200x200 workload; six waves spill in hot paths and run slower.
template <int MODE, int NT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(NT <= 4 ? 5 : 1, 8))) void k_search(
    SearchArgs a) {
  extern __shared__ uint32_t lds_rings[];
  const uint32_t lane = lane_id();
  const uint32_t wib = uniform(threadIdx.x / kWave);
  const uint32_t wid = uniform(blockIdx.x * (blockDim.x / kWave) + wib);
  const uint32_t xcd = a.xsplit ? blockIdx.x % kXcd : 0u;
  const WaveWs w = wave_ws(a, wid, lds_rings + (size_t)wib * (a.nbk_lds + kWaveLds));
  for (uint32_t r = lane; r < a.nbk_lds; r += kWave) lds_st(&w.ring[r], 0u);
  wave_fence();
  uint32_t epoch = a.epochs[wid];
  // sources this wave has finished in the call so far (earlier launches included)
  uint32_t pdone = a.progress ? uniform(__hip_atomic_load(&a.progress[wid], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_SYSTEM)) : 0u;
  unsigned long long st_reached = 0, st_within = 0;  // wave-uniform
  uint32_t st_edges = 0;                             // per lane
  for (;;) {
    const uint32_t item = next_item(a, xcd);
    if (item >= a.n_items) break;
    uint32_t job = 0, si = item;
    if (a.items) {
      const uint2 it = a.items[item];
      job = uniform(it.x);
      si = uniform(it.y);
    }
    CJob &j = ((CJob *)a.jobs)[job];
    const float speed = j.speed;
    SrcCtx c;
    c.src = j.sources[si];
    c.wt = (MODE == kBrandes) ? j.src_wt[si] : 0.0f;
    if (epoch >= kEpochMax) {
      // the 9-bit tags are used up: clear this wave's node records and start over
      for (size_t i = lane; i < a.ncap; i += kWave) w.node[i] = 0ull;
      wave_fence();
      epoch = 0;
    }
    ++epoch;
    c.tag = epoch;
    uint32_t nband[NT], eband[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) nband[t] = eband[t] = 0;
    bool ok = search_source<MODE, NT>(a, j, w, c, nband, eband, st_edges);
    if constexpr (MODE == kBrandes || MODE == kBrandesExport) {
      if (ok && (j.tol > kTieEps || (j.flags & 16u)) && (MODE == kBrandesExport || (j.flags & 2u))) {
        uint32_t pfail = 0;
        ok = j.tol > kTieEps ? phase2_rebuild(j, w, c, pfail) : phase1_fold_after(j, w, c, pfail);
        if (!ok) c.fail |= pfail ? pfail : kFailInternal;
      }
    }
    if (!ok) {
      if (lane == 0) {
        const uint32_t q = atomicAdd(a.n_fail, 1u);
        a.fail_list[q] = item;
        atomicOr(a.fail_flags, c.fail);
      }
      continue;
    }
    const uint32_t R = c.nslot;
    st_reached += R;

    if constexpr (MODE == kBrandes) {
      const float wt = j.src_wt[si];
      const float cyc = j.src_cyc[si];
      float rank[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const uint32_t nn = nband[t];
        const uint32_t ee = eband[t];
        const int r = (int)ee - (int)nn + 1;
        rank[t] = (nn == 0) ? 0.0f : (float)(r > 0 ? r : 0);
        // targets within the threshold: every node counted but the source
        const uint32_t tot = nn ? nn - 1 : 0u;  // padded thresholds hold nobody
        st_within += tot;
        if ((j.flags & 5u) == 5u && lane == 0 && tot) atomicAdd(&j.reach[t], tot);
      }
      // ---- OD-weighted seeds (centrality.rs:2500-2509): the source's destinations
      // that were reached take their trip weight; every other slot keeps 0
      const bool od = (j.flags & 8u) != 0;
      if (od) {
        for (uint32_t k = j.od_off[si] + lane; k < j.od_off[si + 1]; k += kWave) {
          const uint64_t kd = ld64(nkey(w, j.od_dest[k]));
          if (key_current(kd, c.tag)) std_(&w.ods[key_slot(kd)], (double)j.od_w[k]);
        }
        wave_fence();
      }
      // ---- betweenness dependencies: reverse bucket order (centrality.rs:1783-1835,
      // 823-873); each slot's accumulated delta is credited by the flush below
      if (j.flags & 2u) {
        // with Phase 2, a successor may share its predecessor's bucket: visit
        // each bucket level by level, highest first (levels are all 0 otherwise)
        const bool p2 = j.tol > kTieEps || (j.flags & 16u);
        uint32_t s1 = R;
        uint32_t s0 = uniform(ld32(&w.bnd[c.nb - 1]));
        for (int b = (int)c.nb - 1; b >= 0; --b) {
          const uint32_t s0n = (b > 0) ? ld32(&w.bnd[b - 1]) : 0u;
          const uint32_t top = p2 ? uniform(ld32(&w.lvl[b])) : 0u;
          // thresholds some node of this bucket may lie within: a route here is at
          // least (bucket - 1) bucket widths, so their dependencies are loaded
          // with the slot records instead of after them
          bool maybe[NT];
          {
            const uint32_t kb = uniform(ld32(&w.bk[b]));
            const double lo_route = (double)(kb > 0 ? kb - 1 : 0) / (double)j.inv_delta * (double)speed;
#pragma unroll
            for (int t = 0; t < NT; ++t) maybe[t] = !p2 && lo_route <= (double)j.dist[t];
          }
          for (int L = (int)top; L >= 0; --L) {
            // the source (slot 0) has no predecessors
            for (uint32_t s = (s0 ? s0 : 1u) + lane; s < s1; s += kWave) {
              const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
              const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
              double d0[NT], d1[NT];
#pragma unroll
              for (int t = 0; t < NT; ++t) {
                d0[t] = maybe[t] ? ldd(dep_at(w, t, s)) : 0.0;
                d1[t] = maybe[t] ? ldd(dep_at(w, NT + t, s)) : 0.0;
              }
              if (p2 && (uint32_t)(hi >> 48) != (uint32_t)L) continue;
              const uint32_t v = (uint32_t)lo;
              const float route = __uint_as_float((uint32_t)(lo >> 32)) * speed;
              const uint32_t pslot0 = (uint32_t)hi;
              const uint32_t meta = (uint32_t)(hi >> 32);
              const uint32_t npred = meta & 0xffu;
              const double pc = od ? ldd(&w.ods[s]) : ((meta & 0x100u) ? 0.5 : 1.0);
              double dep[NT], dep_b[NT];
              bool any = false;
#pragma unroll
              for (int t = 0; t < NT; ++t) {
                dep[t] = dep_b[t] = 0.0;
                if (route <= j.dist[t]) {
                  const double seed_b = pc * exp(-(double)j.beta[t] * (double)route);
                  dep[t] = pc + (maybe[t] ? d0[t] : ldd(dep_at(w, t, s)));
                  dep_b[t] = seed_b + (maybe[t] ? d1[t] : ldd(dep_at(w, NT + t, s)));
                  any = any || dep[t] != 0.0 || dep_b[t] != 0.0;
                }
              }
              if (!any) continue;
              if (npred == 1) {
                // a single predecessor has sigma_u == sigma_w (sigma is the sum over
                // the predecessors), so sigma_u / sigma_w is exactly 1
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                  if (dep[t] != 0.0 || dep_b[t] != 0.0) {
                    wadd_f64(dep_at(w, t, pslot0), dep[t]);
                    wadd_f64(dep_at(w, NT + t, pslot0), dep_b[t]);
                  }
                }
              } else if (npred > 1) {
                const uint64_t pm = ld64(&w.pm[s]);
                const uint4 rv = j.rows[v];
                const uint32_t f0 = rv.x;
                const bool wide = (rv.y - rv.x) > kWideRow;
                const double sigma_w = ldd(&w.sig[s]);
                uint32_t jb;
                for (uint64_t mm = pm; ps_next(mm, wide, jb);) {
                  const uint32_t u = j.fwd_adj[f0 + jb].x & kNodeMask;
                  const uint32_t su = key_slot(ld64(nkey(w, u)));
                  const double sigma_u = ldd(&w.sig[su]);
                  if (sigma_u == 0.0) continue;
                  const double fct = sigma_u / sigma_w;
#pragma unroll
                  for (int t = 0; t < NT; ++t) {
                    if (dep[t] != 0.0 || dep_b[t] != 0.0) {
                      wadd_f64(dep_at(w, t, su), fct * dep[t]);
                      wadd_f64(dep_at(w, NT + t, su), fct * dep_b[t]);
                    }
                  }
                }
              }
            }
            wave_fence();
          }
          s1 = s0;
          s0 = uniform(s0n);
        }
      }
      // ---- every target's closeness terms and betweenness credit, in node order
      flush_targets<NT>(j, w, c, R, rank, wt, cyc);
    } else if constexpr (MODE == kSegment) {
#pragma unroll
      for (int t = 0; t < NT; ++t) st_within += nband[t] ? nband[t] - 1 : 0u;
      // ---- closeness at the source over visited edges (centrality.rs:2201-2317)
      if (j.flags & 1u) {
        double sd[NT], sh[NT], sb[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) sd[t] = sh[t] = sb[t] = 0.0;
        for (uint32_t s = lane; s < R; s += kWave) {
          const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
          const uint32_t v = (uint32_t)lo;
          const float secs_v = __uint_as_float((uint32_t)(lo >> 32));
          const float sd_n = secs_v * speed;
          const uint4 rw = j.rows[v];
          const uint32_t r0 = rw.z, nr = rw.w - rw.z;
          for (uint32_t cb = 0; cb < nr; cb += kChunk) {
            uint2 re[kChunk];
            float2 tw[kChunk];
            uint64_t kn[kChunk];
#pragma unroll
            for (int q = 0; q < kChunk; ++q) {
              const bool in = cb + q < nr;
              re[q] = in ? j.rev_adj[r0 + cb + q] : make_uint2(v, 0u);
              tw[q] = in ? j.rev_tw[r0 + cb + q] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int q = 0; q < kChunk; ++q) {
              const uint32_t nb = re[q].x & kNodeMask;
              kn[q] = (nb != v) ? ld64(nkey(w, nb)) : 0;
            }
#pragma unroll
            for (int q = 0; q < kChunk; ++q) {
              if (cb + q >= nr) continue;
              const uint32_t nb = re[q].x & kNodeMask;
              float sd_m = INFINITY;
              bool visited_edge = true;
              if (nb == v) {
                sd_m = sd_n;
              } else if (key_current(kn[q], c.tag)) {
                const float s_nb = key_secs(kn[q]);
                visited_edge = key_less(j.orig, secs_v, v, s_nb, nb);
                sd_m = s_nb * speed;
              }
              if (!visited_edge) continue;
              const bool n_nearer = sd_n <= sd_m;
              const float av = n_nearer ? sd_n : sd_m;
              const float bv = n_nearer ? sd_m : sd_n;
              const float a_imp = av, b_imp = bv;
              const float elen = tw[q].x, imp = tw[q].y;
              const float cc = (elen + av + bv) / 2.0f;
              const float dd = cc;
              const float c_imp = a_imp + (cc - av) * imp;
              const float d_imp = c_imp;
#pragma unroll
              for (int t = 0; t < NT; ++t) {
                const float dist = j.dist[t];
                const float beta = j.beta[t];
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
                  const float bb = beta == 0.0f ? curi - a_imp : (exp_f32(nbeta * curi) - exp_f32(nbeta * a_imp)) * inv;
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
                  const float bb = beta == 0.0f ? curi - b_imp : (exp_f32(nbeta * curi) - exp_f32(nbeta * b_imp)) * inv;
                  sb[t] += (double)bb;
                }
              }
            }
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const double t0 = wave_sum(sd[t]), t1 = wave_sum(sh[t]), t2 = wave_sum(sb[t]);
          if (lane == 0) {
            const uint32_t ts = j.tstride;
            const size_t nacc = j.nacc;
            double *accs = j.acc + c.src;
            gadd_f64(&accs[((size_t)0 * ts + t) * nacc], t0);
            gadd_f64(&accs[((size_t)1 * ts + t) * nacc], t1);
            gadd_f64(&accs[((size_t)2 * ts + t) * nacc], t2);
          }
        }
      }
      // ---- betweenness: subtree sums along the single-pred tree (centrality.rs:2319-2401)
      if (j.flags & 2u) {
        uint32_t s1 = R;
        uint32_t s0 = uniform(ld32(&w.bnd[c.nb - 1]));
        for (int b = (int)c.nb - 1; b >= 0; --b) {
          const uint32_t s0n = (b > 0) ? ld32(&w.bnd[b - 1]) : 0u;
          for (uint32_t s = s0 + lane; s < s1; s += kWave) {
            if (s == 0) continue;
            const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
            const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
            const uint64_t lb = ld64(&w.pm[s]);
            double below[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t) below[t] = ldd(&w.acc[(size_t)s * NT + t]);
            const uint32_t v = (uint32_t)lo;
            const float sdist = __uint_as_float((uint32_t)(lo >> 32)) * speed;
            const uint32_t pslot = (uint32_t)hi;
            const float llen = __uint_as_float((uint32_t)lb);
            const float olen = __uint_as_float((uint32_t)(lb >> 32));
            const float min_span = sdist - olen - llen;
            const float o_1 = min_span, o_2 = min_span + olen, l_1 = min_span, l_2 = min_span + llen;
#pragma unroll
            for (int t = 0; t < NT; ++t) {
              if (below[t] > 0.0) gadd_f64(&j.acc[((size_t)3 * j.tstride + t) * j.nacc + v], below[t]);
              double own = 0.0;
              if (j.orig[v] > j.orig[c.src]) {  // pairs counted once, in original index order
                const float dist = j.dist[t];
                const float beta = j.beta[t];
                if (min_span <= dist) {
                  const float o2s = fminf(o_2, dist), l2s = fminf(l_2, dist);
                  float auc;
                  if (beta == 0.0f) {
                    auc = (o2s - o_1) + (l2s - l_1);
                  } else {
                    const float nb2 = -beta, inv = 1.0f / nb2;
                    auc = (exp_f32(nb2 * o2s) - exp_f32(nb2 * o_1)) * inv + (exp_f32(nb2 * l2s) - exp_f32(nb2 * l_1)) * inv;
                  }
                  if (isfinite(auc) && auc >= 0.0f) own = (double)auc;
                }
              }
              const double up = below[t] + own;
              if (pslot != 0 && up != 0.0) wadd_f64(&w.acc[(size_t)pslot * NT + t], up);
            }
          }
          wave_fence();
          s1 = s0;
          s0 = uniform(s0n);
        }
      }
    } else if constexpr (MODE == kAccess) {
      // group g owns records nacc + 2g and nacc + 2g + 1 (node keys to earlier
      // sources): whatever those left there, keys of a larger graph included,
      // carries an older epoch and compares below the current tag
      uint64_t *grec = w.node + (size_t)j.nacc;
      char *stb = a.ws + (size_t)wid * a.ws_stride + a.o_st;
      const auto reach = [&](uint32_t s) { return ld64(&w.rec[2 * (size_t)s]); };
      if (!access_fold(a.acc, c.src, R, reach, j.maxsec, speed, grec, (uint64_t)c.tag << kEpochShift, stb,
                       a.st_cap)) {
        // more winners than the list holds: the source runs again with longer lists
        if (lane == 0) {
          const uint32_t q = atomicAdd(a.n_fail, 1u);
          a.fail_list[q] = item;
          atomicOr(a.fail_flags, kFailCap);
        }
        continue;
      }
    } else {
      // single-source exports (trees / Brandes state)
      for (uint32_t s = lane; s < R; s += kWave) {
        const uint64_t lo = ld64(&w.rec[2 * (size_t)s]);
        const uint32_t v = (uint32_t)lo;
        a.tout.secs[v] = __uint_as_float((uint32_t)(lo >> 32));
        if constexpr (MODE == kTree) {
          const uint64_t hi = ld64(&w.rec[2 * (size_t)s + 1]);
          if (s != 0) {
            const uint32_t pslot = (uint32_t)hi;
            a.tout.pred[v] = (int32_t)(uint32_t)ld64(&w.rec[2 * (size_t)pslot]);
            a.tout.last_eid[v] = (int32_t)j.fwd_eid[j.rows[v].x + (uint32_t)(hi >> 32)];
          }
        } else if constexpr (MODE == kBrandesExport) {
          const uint64_t pm = ld64(&w.pm[s]);
          a.tout.sigma[v] = ldd(&w.sig[s]);
          uint32_t np = 0;
          const uint4 rv = j.rows[v];
          const uint32_t f0 = rv.x;
          const bool wide = (rv.y - rv.x) > kWideRow;
          uint32_t jb;
          for (uint64_t mm = pm; ps_next(mm, wide, jb);) {
            if (np < a.tout.max_preds)
              a.tout.pred_list[(size_t)v * a.tout.max_preds + np] = j.fwd_adj[f0 + jb].x & kNodeMask;
            ++np;
          }
          a.tout.n_preds[v] = np;
        }
      }
      wave_fence();
    }
    if (a.progress) {
      ++pdone;
      if (lane == 0) __hip_atomic_store(&a.progress[wid], pdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  if (lane == 0) {
    a.epochs[wid] = epoch;
    atomicAdd(&a.stats[0], st_reached);
    atomicAdd(&a.stats[2], st_within);
  }
  const unsigned long long edges = wave_sum((unsigned long long)st_edges);
  if (lane == 0) atomicAdd(&a.stats[1], edges);
}

// ---------------------------------------------------------------------------
// Host side: job tables, workspace sizing and launch with retry.
// ---------------------------------------------------------------------------

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

[[noreturn]] void fail_wide_row() {
  fail(CS_EINVAL, "a node with more than 64 out-edges has more than 4 tied predecessors in one search, which the "
                  "GPU search does not support");
}

template <int MODE, int NT>
void launch_one(uint32_t blocks, uint32_t block, size_t lds, hipStream_t stream, const SearchArgs &a) {
  hipLaunchKernelGGL((k_search<MODE, NT>), dim3(blocks), dim3(block), lds, stream, a);
}

template <int MODE, int NT>
int occupancy_one(uint32_t block, size_t lds) {
  int nb = 0;
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_search<MODE, NT>, (int)block, lds),
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  return nb;
}

#define CS_NT_SWITCH(T, CALL)                          \
  switch (T) {                                         \
    case 1: CALL(1); break;                            \
    case 2: CALL(2); break;                            \
    case 3: CALL(3); break;                            \
    case 4: CALL(4); break;                            \
    case 5: CALL(5); break;                            \
    case 6: CALL(6); break;                            \
    case 7: CALL(7); break;                            \
    case 8: CALL(8); break;                            \
    default: fail(CS_EINVAL, "unsupported number of thresholds"); \
  }

template <int MODE>
void launch_nt(uint32_t T, uint32_t blocks, uint32_t block, size_t lds, hipStream_t stream, const SearchArgs &a) {
  if constexpr (MODE == kTree || MODE == kBrandesExport || MODE == kAccess) {
    launch_one<MODE, 1>(blocks, block, lds, stream, a);
  } else {
#define CS_CALL(n) launch_one<MODE, n>(blocks, block, lds, stream, a)
    CS_NT_SWITCH(T, CS_CALL)
#undef CS_CALL
  }
}

template <int MODE>
int blocks_per_cu(uint32_t T, uint32_t block, size_t lds) {
  int nb = 0;
  if constexpr (MODE == kTree || MODE == kBrandesExport || MODE == kAccess) {
    nb = occupancy_one<MODE, 1>(block, lds);
  } else {
#define CS_CALL(n) nb = occupancy_one<MODE, n>(block, lds)
    CS_NT_SWITCH(T, CS_CALL)
#undef CS_CALL
  }
  return nb;
}

struct Plan {
  uint32_t bcap, cap, maxb, nbk, waves;
  size_t stride;
  size_t o_rec, o_pm, o_acc, o_ods, o_sig, o_bkt, o_bnd, o_lvl, o_bk;
  size_t o_st = 0;
  uint32_t st_cap = 0;  // winners per source for land-use statistics (0: none)
};

// Per-wave scratch layout. Everything here is written before it is read
// within one source, so the scratch is never cleared; only the keys (a
// separate arena) carry state across sources.
void layout(Plan &p, uint32_t N, uint32_t T, bool brandes) {
  size_t o = 0;
  p.o_rec = o;
  o = align256(o + 16ull * p.cap);
  p.o_pm = o;
  o = align256(o + 8ull * p.cap);
  p.o_acc = o;
  o = align256(o + 8ull * p.cap * (brandes ? 2 * T : T));
  p.o_ods = o;
  o = align256(o + (brandes ? 8ull * p.cap : 0ull));
  p.o_sig = o;
  o = align256(o + (brandes ? 8ull * p.cap : 0ull));
  p.o_bkt = o;
  o = align256(o + 8ull * p.nbk * p.bcap);
  p.o_bnd = o;
  o = align256(o + 4ull * (p.maxb + 2));
  p.o_lvl = o;
  o = align256(o + 4ull * (p.maxb + 2));
  p.o_bk = o;
  o = align256(o + 4ull * (p.maxb + 2));
  p.o_st = o;
  // winner count | winners, 2 cap | sort keys, 2 cap | values | weights (sorts pad to a power of two)
  if (p.st_cap) o = align256(o + 256 + 16ull * p.st_cap + 16ull * p.st_cap + 8ull * p.st_cap);
  p.stride = o;
}

// Workgroup size for a launch whose widest ring has nbk buckets: the tuned
// size, lowered until the waves' LDS (ring counters plus flush and relaxation
// lists) fits one workgroup's 64 KiB.
uint32_t launch_block(const Tuning &tune, uint32_t nbk) {
  uint32_t block = tune.block ? tune.block : kBlock;
  if (block % kWave || block > (uint32_t)kBlock) fail(CS_EINVAL, "workgroup size must be 64, 128, 192 or 256");
  const uint32_t max_wpb = std::max(1u, kLdsWords / (nbk + kWaveLds));
  return std::min(block, max_wpb * kWave);
}

// Bucket width: below the smallest edge cost by the largest tie-band drift a
// node's fold can accumulate (one 1e-4 step per candidate edge), so candidates
// from the same bucket can never reach a receiver's tie band. When the
// threshold is too long against the shortest edge for such a width to exist,
// the width is 0.95 of the shortest edge (distances stay exact: no node
// improves another of its own bucket) and exact_fold asks for the fold after
// the search, which takes same-bucket candidates in settle order.
float choose_delta(const Graph &g, const DevEdges &de, float maxsec, bool &exact_fold) {
  exact_fold = false;
  const float min_es = de.min_es;
  if (std::isinf(min_es)) return 2.0f * std::max(maxsec, 1.0f);  // no edge between distinct nodes: one bucket
  if (!(min_es > 0.0f)) fail(CS_EINVAL, "graph has zero-cost edges; the bucketed search needs positive edge costs");
  const double deg = (double)std::max(g.max_out_deg, g.max_in_deg);
  const double gap = (double)maxsec * 1.02e-4 * (deg + 3.0) + 1e-3;
  const double delta = (double)min_es - gap;
  if (delta < 0.05 * (double)min_es) {
    exact_fold = true;
    return 0.95f * min_es;
  }
  return (float)delta;
}

// Grow-only device buffer.
struct DevVec {
  void *p = nullptr;
  size_t bytes = 0;
  void *get(size_t need, const char *what) {
    if (need > bytes) {
      if (p) hip_check(hipFree(p), "hipFree");
      p = nullptr;
      bytes = 0;
      if (hipMalloc(&p, need) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        fail(CS_ENOMEM, std::string("could not allocate ") + std::to_string(need >> 20) + " MiB of " + what);
      }
      bytes = need;
    }
    return p;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct Workspace {
  int device = -1;
  // node records (epoch-tagged keys), [key_waves][ncap]: zeroed whenever
  // (re)allocated; each wave clears its own when its 9-bit epochs wrap, and
  // keys left by earlier graphs carry older epochs
  uint64_t *keys = nullptr;
  uint32_t ncap = 0, key_waves = 0;
  uint32_t *epochs = nullptr;
  uint64_t epoch_hi = 0;  // upper bound on epochs used
  DevVec scratch;         // [waves][stride], never cleared
  DevVec jobs, items, ctl, stats, fail;  // launch control
  std::mutex mu;
  void release() {
    if (keys) (void)hipFree(keys);
    if (epochs) (void)hipFree(epochs);
    keys = nullptr;
    epochs = nullptr;
    ncap = key_waves = 0;
    epoch_hi = 0;
    for (DevVec *d : {&scratch, &jobs, &items, &ctl, &stats, &fail}) d->release();
  }
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

// Device bytes the workspace still has to allocate for this shape.
size_t ws_growth(const Workspace &w, uint32_t N, size_t stride, uint32_t waves) {
  size_t g = 0;
  if (N > w.ncap || waves > w.key_waves)
    g += kRecBytes * std::max(N, w.ncap) * std::max(waves, w.key_waves);
  if (stride * waves > w.scratch.bytes) g += stride * waves;
  return g;
}

// Shape the workspace for `waves` waves over N nodes with `stride` bytes of
// scratch each, reserving `sources` epochs.
void prepare_ws(Workspace &w, uint32_t N, size_t stride, uint32_t waves, uint64_t sources,
                hipStream_t stream) {
  bool clear = false;
  if (N > w.ncap || waves > w.key_waves) {
    const uint32_t ncap = std::max(N, w.ncap), kw = std::max(waves, w.key_waves);
    if (w.keys) hip_check(hipFree(w.keys), "hipFree");
    if (w.epochs) hip_check(hipFree(w.epochs), "hipFree");
    w.keys = nullptr;
    w.epochs = nullptr;
    w.ncap = w.key_waves = 0;
    if (hipMalloc(&w.keys, kRecBytes * ncap * kw) != hipSuccess) {
      (void)hipGetLastError();
      w.keys = nullptr;
      fail(CS_ENOMEM, "could not allocate " + std::to_string((kRecBytes * ncap * kw) >> 20) + " MiB of node records");
    }
    hip_check(hipMalloc(&w.epochs, 4ull * kw), "hipMalloc epochs");
    w.ncap = ncap;
    w.key_waves = kw;
    clear = true;
  }
  if (w.epoch_hi + sources + 2 >= 0xfffffff0ull) clear = true;
  if (clear) {
    // every key carries epoch 0, which is never issued
    hip_check(hipMemsetAsync(w.keys, 0, kRecBytes * w.ncap * w.key_waves, stream), "hipMemset keys");
    hip_check(hipMemsetAsync(w.epochs, 0, 4ull * w.key_waves, stream), "hipMemset epochs");
    w.epoch_hi = 0;
  }
  w.scratch.get(stride * waves, "search workspace");
  w.epoch_hi += sources;
}

struct DevInfo {
  uint32_t cus = 0;
};

const DevInfo &dev_info(int dev) {
  static std::mutex mu;
  static std::map<int, DevInfo> cache;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  DevInfo d;
  d.cus = (uint32_t)prop.multiProcessorCount;
  return cache[dev] = d;
}

template <typename T>
struct DevArr {
  T *p = nullptr;
  explicit DevArr(size_t n) {
    if (n) hip_check(hipMalloc(&p, sizeof(T) * n), "hipMalloc");
  }
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  DevArr(const DevArr &) = delete;
  DevArr &operator=(const DevArr &) = delete;
};

Job make_job(const JobSpec &js, uint32_t NT) {
  Graph &g = *js.g;
  const DevEdges &de = *js.de;
  const CallSpec &cs = js.cs;
  Job j;
  std::memset(&j, 0, sizeof(j));
  bool exact_fold = false;
  const float delta = choose_delta(g, de, cs.maxsec, exact_fold) * js.delta_scale;
  j.rows = js.dg->rows;
  j.orig = js.dg->orig;
  j.rev_adj = de.rev_adj;
  j.fwd_adj = de.fwd_adj;
  j.rev_tw = de.rev_tw;
  j.fwd_twlen = de.fwd_twlen;
  j.fwd_eid = de.fwd_eid;
  j.eligible = js.d_eligible;
  j.sources = js.d_sources;
  j.src_wt = js.d_wt;
  j.src_cyc = js.d_cyc;
  j.acc = js.d_acc;
  j.reach = js.d_reach;
  j.maxsec = cs.maxsec;
  j.tol = cs.tolerance;
  j.speed = cs.speed;
  j.inv_delta = 1.0f / delta;
  const uint32_t jump = (uint32_t)std::ceil((double)de.max_es / (double)delta) + 2;
  j.nbk = jump + 1;
  // a wave's ring counters and lists live in LDS, at most 64 KiB per workgroup:
  // wide rings run fewer waves per workgroup (launch_block)
  if (j.nbk > kLdsWords - kWaveLds)
    fail(CS_EINVAL, "edge cost spread too large for the bucket ring (longest edge " + std::to_string(de.max_es) +
                        " s against a bucket width of " + std::to_string(delta) + " s)");
  j.maxb = (uint32_t)((double)cs.maxsec * (double)j.inv_delta) + 4;
  j.flags = (cs.closeness ? 1u : 0u) | (cs.betweenness ? 2u : 0u) | (cs.track_reach ? 4u : 0u) |
            (exact_fold ? 16u : 0u) | (de.has_out_canon ? 32u : 0u);
  // the device's far_factor(max out-degree), the same products in the same order
  j.far_max = 1.0;
  for (uint32_t i = 0; i < g.max_out_deg + 3; ++i) j.far_max *= 1.00010002;
  j.od_off = js.d_od_off;
  j.od_dest = js.d_od_dest;
  j.od_w = js.d_od_w;
  if (js.d_od_off) {
    if (cs.closeness || !cs.betweenness) fail(CS_EINTERNAL, "OD-weighted jobs compute betweenness only");
    j.flags |= 8u;
  }
  j.n_sources = js.n_sources;
  j.tstride = js.tstride ? js.tstride : NT;
  j.nacc = g.n;
  if (j.tstride < cs.T) fail(CS_EINTERNAL, "accumulator stride below the threshold count");
  for (uint32_t i = 0; i < kMaxT; ++i) {
    j.dist[i] = i < cs.T ? cs.dist[i] : -1.0f;
    j.beta[i] = i < cs.T ? cs.beta[i] : 0.0f;
  }
  return j;
}

Tuning g_tuning;
std::mutex g_tuning_mu;

// Runs the sources of every job (or only `explicit_items`, {job, source
// position} pairs) in one persistent launch, retrying sources that outgrew the
// workspace with a larger one.
template <int MODE>
void run_jobs(const std::vector<JobSpec> &specs, uint32_t NT, const Tuning &tune, const TreeOut *tout,
              hipStream_t stream, cs_stats &st, const std::vector<uint2> *explicit_items = nullptr,
              const AccessOut *aout = nullptr, uint32_t *progress = nullptr) {
  constexpr bool kBr = (MODE == kBrandes || MODE == kBrandesExport);
  uint64_t n_items = 0;
  for (auto &js : specs) n_items += js.n_sources;
  if (explicit_items) n_items = explicit_items->size();
  if (n_items == 0) return;
  if (n_items >= 0xffffffffull) fail(CS_EINVAL, "too many sources in one launch");
  const int dev = specs[0].dg->device;
  const DevInfo &di = dev_info(dev);
  std::vector<Job> jobs;
  uint32_t maxN = 0;
  Plan p;
  p.nbk = 1;
  p.maxb = 1;
  for (auto &js : specs) {
    jobs.push_back(make_job(js, NT));
    maxN = std::max(maxN, js.g->n + js.extra_records);
    p.nbk = std::max(p.nbk, jobs.back().nbk);
    p.maxb = std::max(p.maxb, jobs.back().maxb);
  }
  const uint32_t block = launch_block(tune, p.nbk);
  const uint32_t wpb = block / kWave;
  p.bcap = 1024;
  p.cap = std::min<uint32_t>({maxN, tune.cap ? tune.cap : 16384u, kSlotLimit});
  if (aout && aout->stats) p.st_cap = std::max<uint32_t>(1u, std::min<uint32_t>(aout->n_groups, 4096u));
  const size_t lds = (size_t)wpb * (p.nbk + kWaveLds) * 4;
  const int bpc = std::max(1, blocks_per_cu<MODE>(NT, block, lds));
  uint32_t resident = di.cus * (uint32_t)bpc * wpb;
  if (tune.waves_per_simd) resident = std::min<uint32_t>(resident, di.cus * 4 * tune.waves_per_simd);

  Workspace *wsp = acquire_ws(dev);
  WsGuard guard{wsp};
  Job *d_jobs = static_cast<Job *>(wsp->jobs.get(sizeof(Job) * jobs.size(), "job table"));
  hip_check(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(Job) * jobs.size(), hipMemcpyHostToDevice, stream),
            "copy jobs");
  std::vector<uint2> items;  // explicit items only for several jobs or a retry pass
  if (explicit_items) {
    items = *explicit_items;
  } else if (specs.size() > 1) {
    items.reserve(n_items);
    for (uint32_t jb = 0; jb < specs.size(); ++jb)
      for (uint32_t i = 0; i < specs[jb].n_sources; ++i) items.push_back(make_uint2(jb, i));
  }
  uint2 *d_items = static_cast<uint2 *>(wsp->items.get(sizeof(uint2) * std::max<size_t>(items.size(), 1), "items"));
  if (!items.empty())
    hip_check(hipMemcpyAsync(d_items, items.data(), sizeof(uint2) * items.size(), hipMemcpyHostToDevice, stream),
              "copy items");
  const bool use_items = !items.empty();
  uint32_t *d_ctl = static_cast<uint32_t *>(wsp->ctl.get(64, "control words"));
  unsigned long long *d_stats = static_cast<unsigned long long *>(wsp->stats.get(32, "statistics"));
  uint32_t *d_fail = static_cast<uint32_t *>(wsp->fail.get(4 * n_items, "failure list"));
  hip_check(hipMemsetAsync(d_stats, 0, 32, stream), "memset stats");

  uint32_t n_list = (uint32_t)n_items;
  bool list_items = use_items;
  hipEvent_t e0, e1;
  hip_check(hipEventCreate(&e0), "event");
  hip_check(hipEventCreate(&e1), "event");
  double kms = 0.0;
  for (int attempt = 0; attempt < 6; ++attempt) {
    layout(p, maxN, NT, kBr);
    // waves: every wave resident at once, bounded by items and, when the
    // workspace must grow, by a quarter of free HBM
    const uint32_t want = std::max<uint32_t>(wpb, (std::min<uint32_t>(n_list, resident) + wpb - 1) / wpb * wpb);
    uint32_t by_mem = want;
    if (ws_growth(*wsp, maxN, p.stride, want) > 0) {
      size_t free_b = 0, total_b = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
      const size_t have = wsp->scratch.bytes + kRecBytes * wsp->ncap * wsp->key_waves;
      const size_t per_wave = p.stride + kRecBytes * maxN;
      by_mem = (uint32_t)std::min<size_t>((free_b / 4 + have) / per_wave, want);
    }
    if (by_mem < wpb) {
      if (p.cap > 4096) {
        p.cap = std::max<uint32_t>(4096, p.cap / 2);
        continue;
      }
      fail(CS_ENOMEM, "graph too large for the search workspace on this device");
    }
    uint32_t waves = std::min(want, by_mem);
    waves = std::max<uint32_t>(wpb, (waves + wpb - 1) / wpb * wpb);
    p.waves = waves;
    prepare_ws(*wsp, maxN, p.stride, waves, n_list, stream);
    hip_check(hipMemsetAsync(d_ctl, 0, 64, stream), "memset ctl");
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.jobs = d_jobs;
    a.items = list_items ? d_items : nullptr;
    a.n_items = n_list;
    a.bcap = p.bcap;
    a.cap = p.cap;
    a.nbk_lds = p.nbk;
    a.stats = d_stats;
    a.keys = wsp->keys;
    a.ncap = wsp->ncap;
    a.ws = static_cast<char *>(wsp->scratch.p);
    a.ws_stride = p.stride;
    a.o_rec = p.o_rec;
    a.o_pm = p.o_pm;
    a.o_acc = p.o_acc;
    a.o_ods = p.o_ods;
    a.o_sig = p.o_sig;
    a.o_bkt = p.o_bkt;
    a.o_bnd = p.o_bnd;
    a.o_lvl = p.o_lvl;
    a.o_bk = p.o_bk;
    a.o_st = p.o_st;
    a.st_cap = p.st_cap;
    a.epochs = wsp->epochs;
    a.next = d_ctl;
    a.xsplit = tune.xcd_split ? 1u : 0u;
    for (uint32_t x = 0; x <= kXcd; ++x)
      a.xbeg[x] = a.xsplit ? (uint32_t)((uint64_t)n_list * x / kXcd) : (x ? n_list : 0u);
    a.n_fail = d_ctl + kXcd;
    a.fail_flags = d_ctl + kXcd + 1;
    a.fail_list = d_fail;
    if (tout) a.tout = *tout;
    if (aout) a.acc = *aout;
    a.progress = waves <= kProgressSlots ? progress : nullptr;
    hip_check(hipEventRecord(e0, stream), "event record");
    launch_nt<MODE>(NT, waves / wpb, block, lds, stream, a);
    hip_check(hipGetLastError(), "k_search launch");
    hip_check(hipEventRecord(e1, stream), "event record");
    uint32_t ctl[16];
    hip_check(hipMemcpyAsync(ctl, d_ctl, 64, hipMemcpyDeviceToHost, stream), "copy ctl");
    hip_check(hipStreamSynchronize(stream), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    kms += ms;
    st.waves = waves;
    const uint32_t n_failed = ctl[kXcd], flags = ctl[kXcd + 1];
    if (n_failed == 0) break;
    if (flags & kFailWide) fail_wide_row();
    if (flags & kFailInternal) fail(CS_EINTERNAL, "search invariant violated (bucket order)");
    if (flags & kFailRing) fail(CS_EINTERNAL, "bucket ring window exceeded");
    st.retries += n_failed;
    // retry the failed items with a larger workspace
    std::vector<uint32_t> fl(n_failed);
    hip_check(hipMemcpy(fl.data(), d_fail, 4ull * n_failed, hipMemcpyDeviceToHost), "copy");
    std::vector<uint2> retry(n_failed);
    for (uint32_t i = 0; i < n_failed; ++i) retry[i] = list_items ? items[fl[i]] : make_uint2(0u, fl[i]);
    items.swap(retry);
    d_items = static_cast<uint2 *>(wsp->items.get(sizeof(uint2) * n_failed, "items"));
    hip_check(hipMemcpy(d_items, items.data(), sizeof(uint2) * n_failed, hipMemcpyHostToDevice), "copy");
    list_items = true;
    n_list = n_failed;
    if ((flags & kFailCap) && p.cap >= kSlotLimit)
      fail(CS_EINVAL, "a source reaches more than 4,194,304 nodes, the search's slot limit");
    if (flags & kFailCap) {
      p.cap = std::min<uint32_t>({maxN, p.cap * 4, kSlotLimit});
      if (p.st_cap) p.st_cap = std::min<uint32_t>(aout->n_groups, p.st_cap * 4);
    }
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
}

Tuning current_tuning() {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  return g_tuning;
}

}  // namespace

// Clears `n` 64-bit words (a kernel rather than a memset node, so that it can
// sit inside a captured hipGraph).
__global__ void k_zero64(uint64_t *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// A fixed set of shortest-path jobs prepared once and launched many times:
// job table, item list, control words and a workspace of its own, captured
// with the clearing of the accumulators into one hipGraph that each run
// replays. Sources that outgrow the workspace are retried outside the graph.
struct ShortestBatch {
  std::vector<JobSpec> specs;
  std::vector<ZeroRegion> zero;
  uint32_t NT = 1;
  int device = 0;
  Tuning tune;
  bool ready = false;
  uint64_t n_items = 0;
  Plan p{};
  uint32_t block = kBlock, wpb = kBlock / kWave, waves = 0;
  size_t lds = 0;
  std::vector<Job> jobs;
  std::vector<uint2> items;
  Job *d_jobs = nullptr;
  uint2 *d_items = nullptr;
  uint32_t *d_ctl = nullptr, *d_fail = nullptr;
  unsigned long long *d_stats = nullptr;
  Workspace ws;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;

  ~ShortestBatch() { release(); }

  void release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    for (void *q : {(void *)d_jobs, (void *)d_items, (void *)d_ctl, (void *)d_fail, (void *)d_stats})
      if (q) (void)hipFree(q);
    d_jobs = nullptr;
    d_items = nullptr;
    d_ctl = d_fail = nullptr;
    d_stats = nullptr;
    ws.release();
    ready = false;
  }

  void prepare(hipStream_t stream) {
    release();
    n_items = 0;
    for (auto &js : specs) n_items += js.n_sources;
    if (n_items >= 0xffffffffull) fail(CS_EINVAL, "too many sources in one batch");
    jobs.clear();
    uint32_t maxN = 0;
    p = Plan{};
    p.nbk = 1;
    p.maxb = 1;
    for (auto &js : specs) {
      JobSpec j2 = js;
      j2.delta_scale = tune.delta_scale;
      jobs.push_back(make_job(j2, NT));
      maxN = std::max(maxN, js.g->n);
      p.nbk = std::max(p.nbk, jobs.back().nbk);
      p.maxb = std::max(p.maxb, jobs.back().maxb);
    }
    block = launch_block(tune, p.nbk);
    wpb = block / kWave;
    p.bcap = 1024;
    p.cap = std::min<uint32_t>({maxN, tune.cap ? tune.cap : 16384u, kSlotLimit});
    lds = (size_t)wpb * (p.nbk + kWaveLds) * 4;
    const DevInfo &di = dev_info(device);
    const int bpc = std::max(1, blocks_per_cu<kBrandes>(NT, block, lds));
    uint32_t resident = di.cus * (uint32_t)bpc * wpb;
    if (tune.waves_per_simd) resident = std::min<uint32_t>(resident, di.cus * 4 * tune.waves_per_simd);
    layout(p, maxN, NT, true);
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const uint32_t by_mem = (uint32_t)std::min<size_t>((free_b / 4) / (p.stride + kRecBytes * maxN), 1u << 20);
    if (by_mem < wpb) fail(CS_ENOMEM, "batch too large for the search workspace on this device");
    waves = std::min<uint32_t>((uint32_t)std::min<uint64_t>(n_items, resident), by_mem);
    waves = std::max<uint32_t>(wpb, (waves + wpb - 1) / wpb * wpb);
    items.clear();
    if (specs.size() > 1) {
      items.reserve(n_items);
      for (uint32_t jb = 0; jb < specs.size(); ++jb)
        for (uint32_t i = 0; i < specs[jb].n_sources; ++i) items.push_back(make_uint2(jb, i));
    }
    hip_check(hipMalloc(&d_jobs, sizeof(Job) * jobs.size()), "hipMalloc jobs");
    hip_check(hipMemcpy(d_jobs, jobs.data(), sizeof(Job) * jobs.size(), hipMemcpyHostToDevice), "copy jobs");
    hip_check(hipMalloc(&d_items, sizeof(uint2) * std::max<uint64_t>(n_items, 1)), "hipMalloc items");
    if (!items.empty())
      hip_check(hipMemcpy(d_items, items.data(), sizeof(uint2) * items.size(), hipMemcpyHostToDevice), "copy");
    hip_check(hipMalloc(&d_ctl, 64), "hipMalloc ctl");
    hip_check(hipMalloc(&d_stats, 32), "hipMalloc stats");
    hip_check(hipMalloc(&d_fail, 4 * std::max<uint64_t>(n_items, 1)), "hipMalloc fail");
    ws.device = device;
    prepare_ws(ws, maxN, p.stride, waves, 0, stream);
    hip_check(hipStreamSynchronize(stream), "sync");
    // capture: clear control words, statistics and accumulators, then search
    SearchArgs a = args();
    hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t *>(d_ctl), (size_t)8);
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t *>(d_stats), (size_t)4);
    for (const ZeroRegion &z : zero) {
      const size_t grid = std::min<size_t>((z.words + 255) / 256, 4096);
      if (z.words) hipLaunchKernelGGL(k_zero64, dim3((uint32_t)grid), dim3(256), 0, stream, z.ptr, z.words);
    }
    launch_nt<kBrandes>(NT, waves / wpb, block, lds, stream, a);
    hipError_t le = hipGetLastError();
    hipError_t ce = hipStreamEndCapture(stream, &graph);
    hip_check(le, "k_search capture");
    hip_check(ce, "end capture");
    hip_check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    ready = true;
  }

  SearchArgs args() const {
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.jobs = d_jobs;
    a.items = items.empty() ? nullptr : d_items;
    a.n_items = (uint32_t)n_items;
    a.bcap = p.bcap;
    a.cap = p.cap;
    a.nbk_lds = p.nbk;
    a.stats = d_stats;
    a.keys = ws.keys;
    a.ncap = ws.ncap;
    a.ws = static_cast<char *>(ws.scratch.p);
    a.ws_stride = p.stride;
    a.o_rec = p.o_rec;
    a.o_pm = p.o_pm;
    a.o_acc = p.o_acc;
    a.o_ods = p.o_ods;
    a.o_sig = p.o_sig;
    a.o_bkt = p.o_bkt;
    a.o_bnd = p.o_bnd;
    a.o_lvl = p.o_lvl;
    a.o_bk = p.o_bk;
    a.epochs = ws.epochs;
    a.next = d_ctl;
    a.xsplit = tune.xcd_split ? 1u : 0u;
    for (uint32_t x = 0; x <= kXcd; ++x)
      a.xbeg[x] = a.xsplit ? (uint32_t)(n_items * x / kXcd) : (x ? (uint32_t)n_items : 0u);
    a.n_fail = d_ctl + kXcd;
    a.fail_flags = d_ctl + kXcd + 1;
    a.fail_list = d_fail;
    return a;
  }

  void run(hipStream_t stream, cs_stats &st) {
    if (!ready) prepare(stream);
    if (n_items == 0) return;
    // epochs: every wave may use up to n_items of them per run
    if (ws.epoch_hi + n_items + 2 >= 0xfffffff0ull) {
      hip_check(hipMemsetAsync(ws.keys, 0, kRecBytes * ws.ncap * ws.key_waves, stream), "hipMemset keys");
      hip_check(hipMemsetAsync(ws.epochs, 0, 4ull * ws.key_waves, stream), "hipMemset epochs");
      ws.epoch_hi = 0;
    }
    ws.epoch_hi += n_items;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream), "event record");
    hip_check(hipGraphLaunch(exec, stream), "graph launch");
    hip_check(hipEventRecord(e1, stream), "event record");
    uint32_t ctl[16];
    unsigned long long stv[4];
    hip_check(hipMemcpyAsync(ctl, d_ctl, 64, hipMemcpyDeviceToHost, stream), "copy ctl");
    hip_check(hipMemcpyAsync(stv, d_stats, 32, hipMemcpyDeviceToHost, stream), "copy stats");
    hip_check(hipStreamSynchronize(stream), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.kernel_ms += ms;
    st.waves = waves;
    st.nodes_reached += stv[0];
    st.edges_scanned += stv[1];
    st.within_sum += stv[2];
    const uint32_t n_failed = ctl[kXcd], flags = ctl[kXcd + 1];
    if (n_failed == 0) return;
    if (flags & kFailWide) fail_wide_row();
    if (flags & kFailInternal) fail(CS_EINTERNAL, "search invariant violated (bucket order)");
    if (flags & kFailRing) fail(CS_EINTERNAL, "bucket ring window exceeded");
    st.retries += n_failed;
    std::vector<uint32_t> fl(n_failed);
    hip_check(hipMemcpy(fl.data(), d_fail, 4ull * n_failed, hipMemcpyDeviceToHost), "copy");
    std::vector<uint2> retry(n_failed);
    for (uint32_t i = 0; i < n_failed; ++i) retry[i] = items.empty() ? make_uint2(0u, fl[i]) : items[fl[i]];
    Tuning t2 = tune;
    if (p.cap >= kSlotLimit) fail(CS_EINVAL, "a source reaches more than 4,194,304 nodes, the search's slot limit");
    t2.cap = std::min<uint32_t>(std::max<uint32_t>(p.cap * 4, 4096), kSlotLimit);
    run_jobs<kBrandes>(specs, NT, t2, nullptr, stream, st, &retry);
  }

  // Times each candidate launch shape on every k-th source and keeps the
  // fastest per source. The jobs' accumulators hold garbage afterwards; the
  // next run clears them.
  Tuning autotune(double fraction, hipStream_t stream) {
    uint64_t total = 0;
    for (auto &js : specs) total += js.n_sources;
    if (total == 0) return tune;
    const uint64_t stride = std::max<uint64_t>(1, (uint64_t)std::llround(1.0 / std::max(fraction, 1e-6)));
    std::vector<uint2> sample;
    uint64_t k = 0;
    for (uint32_t jb = 0; jb < specs.size(); ++jb)
      for (uint32_t i = 0; i < specs[jb].n_sources; ++i, ++k)
        if (k % stride == 0) sample.push_back(make_uint2(jb, i));
    Tuning best = tune;
    double best_ms = INFINITY;
    for (uint32_t blk : {64u, 128u, 256u})
      for (uint32_t wps : {0u, 3u, 2u})
        for (uint32_t xs : {1u, 0u})
        for (float ds : {1.0f, 0.75f}) {
          Tuning t;
          t.block = blk;
          t.waves_per_simd = wps;
          t.delta_scale = ds;
          t.xcd_split = xs;
          std::vector<JobSpec> sp = specs;
          for (auto &j : sp) j.delta_scale = ds;
          cs_stats s1{};
          try {
            run_jobs<kBrandes>(sp, NT, t, nullptr, stream, s1, &sample);
          } catch (const Error &) {
            continue;  // a shape this device cannot run
          }
          if (s1.kernel_ms < best_ms) {
            best_ms = s1.kernel_ms;
            best = t;
          }
        }
    tune = best;
    ready = false;
    return best;
  }
};

void release_search_workspaces() {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  for (auto &dv : g_ws_pool)
    for (auto &w : dv.second) {
      std::lock_guard<std::mutex> wl(w->mu);  // a running call holds it until its launch has ended
      w->release();
    }
}

void set_default_tuning(const Tuning &t) {
  std::lock_guard<std::mutex> lk(g_tuning_mu);
  g_tuning = t;
}

Tuning default_tuning() { return current_tuning(); }

void launch_shortest(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                     const uint32_t *d_sources, const float *d_wt, const float *d_cyc,
                     uint32_t n_sources, const uint8_t *d_eligible, double *d_acc, uint32_t *d_reach,
                     hipStream_t stream, cs_stats &st, uint32_t tstride, uint32_t *progress) {
  const Tuning tune = current_tuning();
  JobSpec js{&g, &dg, &de, cs, d_sources, d_wt, d_cyc, n_sources, d_eligible, d_acc, d_reach, tune.delta_scale};
  js.tstride = tstride;
  run_jobs<kBrandes>({js}, cs.T, tune, nullptr, stream, st, nullptr, nullptr, progress);
}

// Translation between buffers in original node order and the device order.
__global__ void k_map_ids(uint32_t *out, const uint32_t *in, const uint32_t *perm, uint32_t n_ids, uint32_t N) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_ids) {
    const uint32_t v = in[i];
    out[i] = v < N ? perm[v] : v;  // an invalid id stays invalid
  }
}

__global__ void k_gather_u8(uint8_t *out, const uint8_t *in, const uint32_t *orig, uint32_t N) {
  const uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d < N) out[d] = in[orig[d]];
}

// dst[r][orig[d]] += src[r][d] over rows x N: orig is a permutation, so every
// thread owns its destination
__global__ void k_scatter_add_rows(double *dst, const double *src, const uint32_t *orig, uint32_t N, uint32_t rows) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)N * rows) {
    const size_t r = i / N, d = i - r * N;
    dst[r * N + orig[d]] += src[i];
  }
}

// dst[m][t0 + t][orig[d]] += src[m][t][d] (dst rows of stride tstride thresholds, src of T): the
// rows of a threshold chunk scattered back to original node order
__global__ void k_scatter_add_rows_t(double *dst, const double *src, const uint32_t *orig, uint32_t N, uint32_t M,
                                     uint32_t T, uint32_t tstride, uint32_t t0) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < (size_t)N * M * T) {
    const size_t r = i / N, d = i - r * N;
    const size_t m = r / T, t = r - m * T;
    dst[(m * tstride + t0 + t) * N + orig[d]] += src[i];
  }
}

void launch_shortest_original(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                              const uint32_t *d_sources, const float *d_wt, const float *d_cyc,
                              uint32_t n_sources, const uint8_t *d_eligible, double *d_acc, uint32_t *d_reach,
                              hipStream_t stream, cs_stats &st, uint32_t tstride, uint32_t *progress) {
  if (g.identity_order) {
    launch_shortest(g, dg, de, cs, d_sources, d_wt, d_cyc, n_sources, d_eligible, d_acc, d_reach, stream, st,
                    tstride, progress);
    return;
  }
  if (!tstride) tstride = cs.T;
  const uint32_t N = g.n, rows = 7 * cs.T;
  constexpr uint32_t kB = 256;
  const size_t total = (size_t)rows * N;
  // staging, 256-byte aligned parts: acc [rows][N] f64 | sources u32 | eligibility u8
  const size_t acc_b = (total * 8 + 255) & ~(size_t)255, src_b = ((size_t)n_sources * 4 + 255) & ~(size_t)255;
  const size_t need = acc_b + src_b + N;
  if (dg.xbuf_bytes < need) {
    if (dg.xbuf) hip_check(hipFree(dg.xbuf), "hipFree staging");
    dg.xbuf = nullptr;
    dg.xbuf_bytes = 0;
    hip_check(hipMalloc(&dg.xbuf, need), "hipMalloc staging");
    dg.xbuf_bytes = need;
  }
  double *acc = static_cast<double *>(dg.xbuf);
  uint32_t *src = reinterpret_cast<uint32_t *>(static_cast<char *>(dg.xbuf) + acc_b);
  uint8_t *elig = reinterpret_cast<uint8_t *>(static_cast<char *>(dg.xbuf) + acc_b + src_b);
  if (n_sources)
    hipLaunchKernelGGL(k_map_ids, dim3((n_sources + kB - 1) / kB), dim3(kB), 0, stream, src, d_sources, dg.perm,
                       n_sources, N);
  hipLaunchKernelGGL(k_gather_u8, dim3((N + kB - 1) / kB), dim3(kB), 0, stream, elig, d_eligible, dg.orig, N);
  hip_check(hipMemsetAsync(acc, 0, sizeof(double) * total, stream), "memset");
  hip_check(hipGetLastError(), "translate launch");
  launch_shortest(g, dg, de, cs, src, d_wt, d_cyc, n_sources, elig, acc, d_reach, stream, st, 0, progress);
  hipLaunchKernelGGL(k_scatter_add_rows_t, dim3((uint32_t)((total + kB - 1) / kB)), dim3(kB), 0, stream, d_acc, acc,
                     dg.orig, N, 7u, cs.T, tstride, 0u);
  hip_check(hipGetLastError(), "scatter launch");
  // the staging is reused by the graph's next call once the lock is released
  hip_check(hipStreamSynchronize(stream), "sync");
}

// A job table entry written from a kernel argument (no host copy to the device).
__global__ void k_put_job(Job j, Job *dst) {
  if (threadIdx.x == 0) *dst = j;
}

// Items {job 0, source position or earlier item} of a retry pass from a launch's failure list.
__global__ void k_retry_items(uint2 *out, const uint32_t *fail, uint32_t n, const uint2 *prev) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = prev ? prev[fail[i]] : make_uint2(0u, fail[i]);
}

// launch_shortest_original with every device buffer it needs beyond the
// caller's inputs and outputs carved from `ws` (ws_bytes bytes, 256-byte
// aligned): no allocation. The region holds, in order: control words and
// statistics, the job table, the failure list and two retry item lists, the
// node-order staging (graphs with a device order) and then, per wave, its
// epoch, node records and scratch. As many waves run as the rest holds.
void launch_shortest_in(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs, const uint32_t *d_sources,
                        const float *d_wt, const float *d_cyc, uint32_t n_sources, const uint8_t *d_eligible,
                        double *d_acc, uint32_t *d_reach, uint32_t tstride, uint32_t t0, hipStream_t stream,
                        cs_stats &st, char *ws, size_t ws_bytes) {
  if (n_sources == 0) return;
  if (reinterpret_cast<uintptr_t>(ws) & 255u) fail(CS_EINTERNAL, "workspace region not 256-byte aligned");
  const Tuning tune = current_tuning();
  const uint32_t N = g.n, T = cs.T;
  const bool translate = !g.identity_order;
  const DevInfo &di = dev_info(dg.device);
  constexpr uint32_t kB = 256;
  // fixed parts
  size_t o = 0;
  const size_t o_ctl = o;
  o = align256(o + 64);
  const size_t o_stats = o;
  o = align256(o + 32);
  const size_t o_job = o;
  o = align256(o + sizeof(Job));
  const size_t o_fail = o;
  o = align256(o + 4ull * n_sources);
  const size_t o_items0 = o;
  o = align256(o + 8ull * n_sources);
  const size_t o_items1 = o;
  o = align256(o + 8ull * n_sources);
  const size_t rows = 7ull * T;
  size_t o_sacc = 0, o_ssrc = 0, o_selig = 0;
  if (translate) {
    o_sacc = o;
    o = align256(o + 8ull * rows * N);
    o_ssrc = o;
    o = align256(o + 4ull * n_sources);
    o_selig = o;
    o = align256(o + N);
  }
  const size_t fixed = o;
  if (fixed > ws_bytes) fail(CS_ENOMEM, "call too large for the allocation-free workspace");
  uint32_t *d_ctl = reinterpret_cast<uint32_t *>(ws + o_ctl);
  unsigned long long *d_stats = reinterpret_cast<unsigned long long *>(ws + o_stats);
  Job *d_job = reinterpret_cast<Job *>(ws + o_job);
  uint32_t *d_fail = reinterpret_cast<uint32_t *>(ws + o_fail);
  uint2 *d_items[2] = {reinterpret_cast<uint2 *>(ws + o_items0), reinterpret_cast<uint2 *>(ws + o_items1)};

  const uint32_t *src = d_sources;
  const uint8_t *elig = d_eligible;
  double *acc = d_acc + (size_t)t0 * N;
  uint32_t stride_t = tstride;
  if (translate) {
    uint32_t *ssrc = reinterpret_cast<uint32_t *>(ws + o_ssrc);
    uint8_t *selig = reinterpret_cast<uint8_t *>(ws + o_selig);
    acc = reinterpret_cast<double *>(ws + o_sacc);
    stride_t = T;
    hipLaunchKernelGGL(k_map_ids, dim3((n_sources + kB - 1) / kB), dim3(kB), 0, stream, ssrc, d_sources, dg.perm,
                       n_sources, N);
    hipLaunchKernelGGL(k_gather_u8, dim3((N + kB - 1) / kB), dim3(kB), 0, stream, selig, d_eligible, dg.orig, N);
    hipLaunchKernelGGL(k_zero64, dim3(1024), dim3(kB), 0, stream, reinterpret_cast<uint64_t *>(acc), rows * N);
    hip_check(hipGetLastError(), "translate launch");
    src = ssrc;
    elig = selig;
  }
  JobSpec js{&g, &dg, &de, cs, src, d_wt, d_cyc, n_sources, elig, acc, d_reach ? d_reach + t0 : nullptr,
             tune.delta_scale};
  js.tstride = stride_t;
  Job job = make_job(js, T);
  hipLaunchKernelGGL(k_put_job, dim3(1), dim3(64), 0, stream, job, d_job);
  hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t *>(d_stats), (size_t)4);
  hip_check(hipGetLastError(), "job table launch");

  Plan p{};
  p.nbk = job.nbk;
  p.maxb = job.maxb;
  const uint32_t block = launch_block(tune, p.nbk);
  const uint32_t wpb = block / kWave;
  p.bcap = 1024;
  p.cap = std::min<uint32_t>({N, tune.cap ? tune.cap : 16384u, kSlotLimit});
  const size_t lds = (size_t)wpb * (p.nbk + kWaveLds) * 4;
  const int bpc = std::max(1, blocks_per_cu<kBrandes>(T, block, lds));
  uint32_t resident = di.cus * (uint32_t)bpc * wpb;
  if (tune.waves_per_simd) resident = std::min<uint32_t>(resident, di.cus * 4 * tune.waves_per_simd);

  uint32_t n_list = n_sources;
  const uint2 *items = nullptr;
  int cur = 0;
  for (int attempt = 0;; ++attempt) {
    layout(p, N, T, true);
    const size_t per_wave = 4 + kRecBytes * N + p.stride;
    const uint32_t want = std::max<uint32_t>(wpb, (std::min<uint32_t>(n_list, resident) + wpb - 1) / wpb * wpb);
    // epochs, node records and scratch of w waves, each part 256-byte aligned
    auto bytes_for = [&](uint32_t w) { return align256(4ull * w) + align256(kRecBytes * N * w) + p.stride * w; };
    uint32_t waves = std::min<uint32_t>(want, (uint32_t)std::min<size_t>((ws_bytes - fixed) / per_wave, want));
    waves = waves / wpb * wpb;
    while (waves >= wpb && fixed + bytes_for(waves) > ws_bytes) waves -= wpb;
    if (waves < wpb) {
      if (p.cap > 1024) {
        p.cap = std::max<uint32_t>(1024, p.cap / 2);
        continue;
      }
      fail(CS_ENOMEM, "graph of " + std::to_string(N) +
                          " nodes too large for the allocation-free workspace (" +
                          std::to_string(ws_bytes >> 20) + " MiB)");
    }
    char *base = ws + fixed;
    uint32_t *d_epochs = reinterpret_cast<uint32_t *>(base);
    uint64_t *d_keys = reinterpret_cast<uint64_t *>(base + align256(4ull * waves));
    char *d_scratch = base + align256(4ull * waves) + align256(kRecBytes * N * waves);
    // node records and epochs start from zero (epoch 0 is never issued)
    hipLaunchKernelGGL(k_zero64, dim3(2048), dim3(kB), 0, stream, reinterpret_cast<uint64_t *>(base),
                       (align256(4ull * waves) + align256(kRecBytes * N * waves)) / 8);
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(64), 0, stream, reinterpret_cast<uint64_t *>(d_ctl), (size_t)8);
    hip_check(hipGetLastError(), "clear launch");
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.jobs = d_job;
    a.items = items;
    a.n_items = n_list;
    a.bcap = p.bcap;
    a.cap = p.cap;
    a.nbk_lds = p.nbk;
    a.stats = d_stats;
    a.keys = d_keys;
    a.ncap = N;
    a.ws = d_scratch;
    a.ws_stride = p.stride;
    a.o_rec = p.o_rec;
    a.o_pm = p.o_pm;
    a.o_acc = p.o_acc;
    a.o_ods = p.o_ods;
    a.o_sig = p.o_sig;
    a.o_bkt = p.o_bkt;
    a.o_bnd = p.o_bnd;
    a.o_lvl = p.o_lvl;
    a.o_bk = p.o_bk;
    a.epochs = d_epochs;
    a.next = d_ctl;
    a.xsplit = tune.xcd_split ? 1u : 0u;
    for (uint32_t x = 0; x <= kXcd; ++x)
      a.xbeg[x] = a.xsplit ? (uint32_t)((uint64_t)n_list * x / kXcd) : (x ? n_list : 0u);
    a.n_fail = d_ctl + kXcd;
    a.fail_flags = d_ctl + kXcd + 1;
    a.fail_list = d_fail;
    launch_nt<kBrandes>(T, waves / wpb, block, lds, stream, a);
    hip_check(hipGetLastError(), "k_search launch");
    uint32_t ctl[16];
    hip_check(hipMemcpyAsync(ctl, d_ctl, 64, hipMemcpyDeviceToHost, stream), "copy ctl");
    hip_check(hipStreamSynchronize(stream), "sync");
    st.waves = waves;
    const uint32_t n_failed = ctl[kXcd], flags = ctl[kXcd + 1];
    if (n_failed == 0) break;
    if (flags & kFailWide) fail_wide_row();
    if (flags & kFailInternal) fail(CS_EINTERNAL, "search invariant violated (bucket order)");
    if (flags & kFailRing) fail(CS_EINTERNAL, "bucket ring window exceeded");
    if ((flags & kFailCap) && p.cap >= std::min<uint32_t>(N, kSlotLimit))
      fail(CS_EINVAL, "a source reaches more nodes than the search's slot limit");
    if (attempt == 5) fail(CS_EINTERNAL, "search workspace retries exhausted");
    st.retries += n_failed;
    // the failed items, as {job, source position}, run again with more room
    uint2 *next = d_items[cur];
    hipLaunchKernelGGL(k_retry_items, dim3((n_failed + kB - 1) / kB), dim3(kB), 0, stream, next, d_fail, n_failed,
                       items);
    hip_check(hipGetLastError(), "retry items launch");
    items = next;
    cur ^= 1;
    n_list = n_failed;
    if (flags & kFailCap) p.cap = std::min<uint32_t>({N, p.cap * 4, kSlotLimit});
    if (flags & kFailBucket) p.bcap *= 4;
  }
  if (translate) {
    const size_t total = rows * N;
    hipLaunchKernelGGL(k_scatter_add_rows_t, dim3((uint32_t)((total + kB - 1) / kB)), dim3(kB), 0, stream, d_acc, acc,
                       dg.orig, N, 7u, T, tstride, t0);
    hip_check(hipGetLastError(), "scatter launch");
  }
  unsigned long long stv[4];
  hip_check(hipMemcpyAsync(stv, d_stats, 32, hipMemcpyDeviceToHost, stream), "copy stats");
  hip_check(hipStreamSynchronize(stream), "sync");
  st.nodes_reached += stv[0];
  st.edges_scanned += stv[1];
  st.within_sum += stv[2];
}

void launch_shortest_batch(const std::vector<JobSpec> &jobs, uint32_t NT, const Tuning &tune, hipStream_t stream,
                           cs_stats &st, const std::vector<uint2> *items, uint32_t *progress) {
  run_jobs<kBrandes>(jobs, NT, tune, nullptr, stream, st, items, nullptr, progress);
}

ShortestBatch *batch_new(std::vector<JobSpec> specs, std::vector<ZeroRegion> zero, uint32_t NT, int device,
                         const Tuning &tune) {
  auto *b = new ShortestBatch();
  b->specs = std::move(specs);
  b->zero = std::move(zero);
  b->NT = NT;
  b->device = device;
  b->tune = tune;
  return b;
}
void batch_delete(ShortestBatch *b) { delete b; }
void batch_set_tuning(ShortestBatch &b, const Tuning &t) {
  b.tune = t;
  b.ready = false;
}
Tuning batch_tuning(const ShortestBatch &b) { return b.tune; }
Tuning batch_autotune(ShortestBatch &b, double fraction, hipStream_t stream) { return b.autotune(fraction, stream); }
void batch_run(ShortestBatch &b, hipStream_t stream, cs_stats &st) { b.run(stream, st); }

void launch_segment(Graph &g, DevGraph &dg, DevEdges &de, const CallSpec &cs,
                    const uint32_t *d_sources, uint32_t n_sources, double *d_acc, hipStream_t stream,
                    cs_stats &st, uint32_t tstride, uint32_t *progress) {
  const Tuning tune = current_tuning();
  JobSpec js{&g, &dg, &de, cs, d_sources, nullptr, nullptr, n_sources, nullptr, d_acc, nullptr, tune.delta_scale};
  js.tstride = tstride;
  run_jobs<kSegment>({js}, cs.T, tune, nullptr, stream, st, nullptr, nullptr, progress);
}

void launch_access(Graph &g, DevGraph &dg, DevEdges &de, float maxsec, const uint32_t *d_sources,
                   uint32_t n_sources, const AccessOut &out, hipStream_t stream, cs_stats &st,
                   uint32_t *progress) {
  if ((uint64_t)g.n + 2ull * out.n_shared > 0x7fffffffull)
    fail(CS_EINVAL, "too many nodes and dedupe groups for the per-wave records");
  CallSpec cs;
  cs.T = 1;
  cs.dist[0] = maxsec * de.speed;
  cs.maxsec = maxsec;
  cs.speed = de.speed;
  cs.tolerance = kTieEps;
  cs.closeness = false;
  cs.betweenness = false;
  const Tuning tune = current_tuning();
  JobSpec js{&g, &dg, &de, cs, d_sources, nullptr, nullptr, n_sources, nullptr, nullptr, nullptr, tune.delta_scale};
  js.extra_records = 2 * out.n_shared;  // two records (epoch-tagged words) per shared group
  run_jobs<kAccess>({js}, 1, tune, nullptr, stream, st, nullptr, &out, progress);
}

// Data aggregation over reached-node lists that another search produced (the
// simplest-path trees of angular aggregation): one wave per source, persistent
// waves taking sources from a counter. A wave's dedupe-group words start at
// zero (the launch clears them) and carry the wave's own 9-bit epochs, cleared
// again when those run out.
struct ListAccessArgs {
  AccessOut ao;
  const uint32_t *src;     // device node of each source position
  const uint32_t *items;   // [n_run] source position of each item, or null: item i is position i
  const uint32_t *runs;    // [n_run] items to aggregate, or null: 0 .. n_run - 1
  const uint64_t *lists;   // [items][lcap] {node | seconds << 32}
  const uint32_t *counts;  // [items] list length or kNoList
  uint32_t n_run, lcap;
  float maxsec, speed;
  char *ws;
  size_t ws_stride, o_st;
  uint32_t st_cap;
  uint32_t *next, *n_fail, *fail_list;
  uint32_t *progress;  // [waves] sources each wave has finished (pinned host memory), or null
};

__global__ __launch_bounds__(kBlock) void k_access_lists(ListAccessArgs a) {
  const uint32_t lane = lane_id();
  const uint32_t wid = uniform(blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave);
  uint64_t *grec = reinterpret_cast<uint64_t *>(a.ws + (size_t)wid * a.ws_stride);
  char *stb = a.ws + (size_t)wid * a.ws_stride + a.o_st;
  uint32_t pdone = a.progress ? uniform(__hip_atomic_load(&a.progress[wid], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_SYSTEM)) : 0u;
  uint32_t epoch = 0;
  for (;;) {
    uint32_t r = 0;
    if (lane == 0) r = atomicAdd(a.next, 1u);
    r = __shfl(r, 0);
    if (r >= a.n_run) break;
    const uint32_t item = uniform(a.runs ? a.runs[r] : r);
    const uint32_t R = uniform(a.counts[item]);
    if (R == kNoList) continue;
    if (epoch >= kEpochMax) {
      for (size_t i = lane; i < 2 * (size_t)a.ao.n_shared; i += kWave) grec[i] = 0ull;
      wave_fence();
      epoch = 0;
    }
    ++epoch;
    const uint32_t src = uniform(a.src[a.items ? a.items[item] : item]);
    const uint64_t *lst = a.lists + (size_t)item * a.lcap;
    const auto reach = [&](uint32_t s) { return lst[s]; };
    if (!access_fold(a.ao, src, R, reach, a.maxsec, a.speed, grec, (uint64_t)epoch << kEpochShift, stb,
                     a.st_cap)) {
      if (lane == 0) a.fail_list[atomicAdd(a.n_fail, 1u)] = item;
      continue;
    }
    if (a.progress) {
      ++pdone;
      if (lane == 0) __hip_atomic_store(&a.progress[wid], pdone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

void launch_access_lists(const AccessOut &out, const uint32_t *d_src, const uint32_t *d_items,
                         const uint64_t *d_lists, const uint32_t *d_counts, uint32_t n_items, uint32_t lcap,
                         float maxsec, float speed, hipStream_t stream, cs_stats &st, uint32_t *progress) {
  if (n_items == 0) return;
  if (out.n_shared > 0x7fffffffu) fail(CS_EINVAL, "too many dedupe groups for the per-wave records");
  int dev = 0, n_cu = 0;
  hip_check(hipGetDevice(&dev), "device");
  hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
  const size_t wpb = kBlock / kWave;
  size_t waves = std::min<size_t>(n_items, (size_t)n_cu * 16);
  waves = (waves + wpb - 1) / wpb * wpb;
  uint32_t st_cap = out.stats ? std::max<uint32_t>(1u, std::min<uint32_t>(out.n_groups, 4096u)) : 0u;
  std::vector<uint32_t> runs;  // empty: every item
  uint32_t n_run = n_items;
  DevMem<uint32_t> d_ctl(4), d_fail(n_items), d_runs(n_items);
  for (int attempt = 0; attempt < 8; ++attempt) {
    const size_t o_st = align256(16ull * out.n_shared);
    const size_t stride = o_st + (st_cap ? align256(256 + 40ull * st_cap) : 0);
    // per-wave scratch (none without shared groups or statistics) within a
    // quarter of free HBM: fewer waves for long winner lists
    if (stride) {
      size_t free_b = 0, total_b = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "meminfo");
      const size_t fit = (free_b / 4) / stride / wpb * wpb;
      if (fit == 0) fail(CS_ENOMEM, "not enough device memory for the aggregation scratch");
      waves = std::min(waves, fit);
    }
    DevMem<char> ws(waves * stride);
    hip_check(hipMemsetAsync(ws.p, 0, waves * stride, stream), "memset");
    hip_check(hipMemsetAsync(d_ctl.p, 0, 16, stream), "memset");
    if (!runs.empty())
      hip_check(hipMemcpyAsync(d_runs.p, runs.data(), 4ull * n_run, hipMemcpyHostToDevice, stream), "copy");
    ListAccessArgs a;
    a.ao = out;
    a.src = d_src;
    a.items = d_items;
    a.runs = runs.empty() ? nullptr : d_runs.p;
    a.lists = d_lists;
    a.counts = d_counts;
    a.n_run = n_run;
    a.lcap = lcap;
    a.maxsec = maxsec;
    a.speed = speed;
    a.ws = ws.p;
    a.ws_stride = stride;
    a.o_st = o_st;
    a.st_cap = st_cap;
    a.next = d_ctl.p;
    a.n_fail = d_ctl.p + 1;
    a.fail_list = d_fail.p;
    a.progress = waves <= kProgressSlots ? progress : nullptr;
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream), "event");
    hipLaunchKernelGGL(k_access_lists, dim3(waves / wpb), dim3(kBlock), 0, stream, a);
    hip_check(hipGetLastError(), "k_access_lists launch");
    hip_check(hipEventRecord(e1, stream), "event");
    uint32_t ctl[4];
    hip_check(hipMemcpyAsync(ctl, d_ctl.p, 16, hipMemcpyDeviceToHost, stream), "copy");
    hip_check(hipStreamSynchronize(stream), "sync");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    st.kernel_ms += ms;
    if (ctl[1] == 0) return;
    // statistics sources with more winners than the lists hold run again with longer lists
    if (st_cap >= out.n_groups) fail(CS_EINTERNAL, "statistics winner list overflow");
    st.retries += ctl[1];
    runs.resize(ctl[1]);
    hip_check(hipMemcpy(runs.data(), d_fail.p, 4ull * ctl[1], hipMemcpyDeviceToHost), "copy");
    n_run = ctl[1];
    st_cap = std::min<uint32_t>(out.n_groups, st_cap * 4);
  }
  fail(CS_EINTERNAL, "statistics list retries exhausted");
}

// One thread per (threshold, node): the diversity measures of the class
// counts within that threshold (diversity.rs:6-175, 341-385), in f32 with the
// classes summed in index order; a measure the reference reports as an error
// (non-finite) is 0, as mixed_uses' unwrap_or(0.0) makes it (data.rs:866-941).
__global__ void k_diversity(AccessOut ao, uint32_t C, double *div) {
  const uint32_t T = ao.T, N = ao.N, TS = ao.tstride;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)T * N) return;
  const uint32_t t = (uint32_t)(i / N), v = (uint32_t)(i % N);
  float out[kDivMetrics];
  for (uint32_t k = 0; k < kDivMetrics; ++k) out[k] = 0.0f;
  uint32_t num = 0, rich = 0;
  for (uint32_t c = 0; c < C; ++c) {
    const uint32_t n = (uint32_t)ao.count[((size_t)c * TS + t) * N + v];
    num += n;
    rich += n > 0;
  }
  if (num > 0) {
    const float numf = (float)num;
    const float beta = ao.beta[t], mcw = ao.mcw[t];
    float ent = 0.0f, sq = 0.0f, agg = 0.0f, lam = 0.0f;
    const float nm1 = (float)(num - 1);
    for (uint32_t c = 0; c < C; ++c) {
      const uint32_t n = (uint32_t)ao.count[((size_t)c * TS + t) * N + v];
      if (n == 0) continue;
      const float p = __fdiv_rn((float)n, numf);
      ent = __fadd_rn(ent, __fmul_rn(p, logf(p)));
      sq = __fadd_rn(sq, powf(p, 2.0f));
      const float d = (float)__longlong_as_double((long long)ao.nearest[(size_t)c * N + v]);
      const float w = __fdiv_rn(fminf(exp_f32(__fmul_rn(-beta, d)), mcw), mcw);
      agg = __fadd_rn(agg, __fmul_rn(w, p));
      if (num >= 2) lam = __fadd_rn(lam, __fmul_rn(p, __fdiv_rn(fmaxf(__fsub_rn((float)n, 1.0f), 0.0f), nm1)));
    }
    out[0] = (float)rich;
    const float h1 = exp_f32(-ent), h2 = powf(sq, -1.0f);
    out[1] = isfinite(h1) ? h1 : 0.0f;
    out[2] = isfinite(sq) && isfinite(h2) ? h2 : 0.0f;
    if (!(fabsf(agg) < __FLT_EPSILON__)) {
      float s0 = 0.0f, we = 0.0f, s2 = 0.0f;
      for (uint32_t c = 0; c < C; ++c) {
        const uint32_t n = (uint32_t)ao.count[((size_t)c * TS + t) * N + v];
        if (n == 0) continue;
        const float p = __fdiv_rn((float)n, numf);
        const float d = (float)__longlong_as_double((long long)ao.nearest[(size_t)c * N + v]);
        const float w = __fdiv_rn(fminf(exp_f32(__fmul_rn(-beta, d)), mcw), mcw);
        const float r = __fdiv_rn(p, agg);
        s0 = __fadd_rn(s0, __fmul_rn(w, powf(r, 0.0f)));
        const float e = __fdiv_rn(__fmul_rn(w, p), agg);
        if (e > 0.0f) we = __fadd_rn(we, __fmul_rn(e, logf(e)));
        s2 = __fadd_rn(s2, __fmul_rn(w, powf(r, 2.0f)));
      }
      const float w0 = powf(s0, 1.0f), w1 = exp_f32(-we), w2 = powf(s2, -1.0f);
      out[3] = isfinite(s0) && isfinite(w0) ? w0 : 0.0f;
      out[4] = isfinite(w1) ? w1 : 0.0f;
      out[5] = isfinite(s2) && isfinite(w2) ? w2 : 0.0f;
    }
    out[6] = isfinite(ent) ? fmaxf(-ent, 0.0f) : 0.0f;
    out[7] = num >= 2 ? fmaxf(__fsub_rn(1.0f, lam), 0.0f) : 0.0f;
  }
  for (uint32_t k = 0; k < kDivMetrics; ++k) div[((size_t)k * TS + t) * N + v] = (double)out[k];
}

void launch_diversity(const AccessOut &out, uint32_t n_classes, double *div, hipStream_t stream) {
  const uint64_t n = (uint64_t)out.T * out.N;
  if (n == 0) return;
  const uint32_t block = 256;
  hipLaunchKernelGGL(k_diversity, dim3((uint32_t)((n + block - 1) / block)), dim3(block), 0, stream, out, n_classes,
                     div);
  hip_check(hipGetLastError(), "k_diversity launch");
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
  DevArr<uint32_t> d_src(1);
  hip_check(hipMemcpy(d_src.p, &src, 4, hipMemcpyHostToDevice), "copy");
  cs_stats st{};
  Tuning tune;  // single-source exports: defaults
  JobSpec js{&g, &dg, &de, cs, d_src.p, nullptr, nullptr, 1, nullptr, nullptr, nullptr, 1.0f};
  if (mode == 0)
    run_jobs<kTree>({js}, 1, tune, &out, stream, st);
  else
    run_jobs<kBrandesExport>({js}, 1, tune, &out, stream, st);
}

}  // namespace cs
