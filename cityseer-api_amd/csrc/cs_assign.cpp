// cs_assign.cpp -- assignment of data geometries to street-network nodes on the
// host (DataMap.assign_data_to_network, rust/src/data.rs:455-512, and
// NetworkStructure.find_assignments_for_entry, rust/src/graph.rs:1346-1535).
//
// Each data geometry takes candidate street edges: the n nearest edge
// envelopes for a point, every edge envelope within the assignment distance of
// the geometry's envelope otherwise. Candidates whose geometry lies within the
// assignment distance are visited nearest first; each one offers its two end
// nodes, a node being valid when the line from the nearest point of the data
// geometry to the node crosses no barrier and no street edge (a proper
// crossing). A point stops at the first edge that assigns a node; other
// geometries visit at most n candidates. Edge envelopes sit in a uniform grid
// instead of an R-tree; the candidate order is the same (envelope distance,
// then true distance), ties going to the edge listed first.
//
// Orientation signs, which decide crossings, are exact: a floating-point
// filter with Shewchuk's error bound, then a 113-bit recomputation in which
// the coordinate differences and products of doubles are exact.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <memory>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "cs_api.hpp"

namespace {

struct P {
  double x, y;
};

// ---- exact orientation -------------------------------------------------------
int orient(const P &a, const P &b, const P &c) {
  const double l = (b.x - a.x) * (c.y - a.y), r = (b.y - a.y) * (c.x - a.x);
  const double det = l - r;
  const double bound = 3.3306690738754716e-16 * (std::fabs(l) + std::fabs(r));
  if (det > bound) return 1;
  if (-det > bound) return -1;
  const __float128 bx = (__float128)b.x - a.x, by = (__float128)b.y - a.y;
  const __float128 cx = (__float128)c.x - a.x, cy = (__float128)c.y - a.y;
  const __float128 d = bx * cy - by * cx;
  return d > 0 ? 1 : (d < 0 ? -1 : 0);
}

bool on_segment_collinear(const P &a, const P &b, const P &p) {
  return std::min(a.x, b.x) <= p.x && p.x <= std::max(a.x, b.x) && std::min(a.y, b.y) <= p.y &&
         p.y <= std::max(a.y, b.y);
}

// segments [a, b] and [c, d] share a point (touching and overlaps included)
bool seg_intersects(const P &a, const P &b, const P &c, const P &d) {
  const int o1 = orient(a, b, c), o2 = orient(a, b, d), o3 = orient(c, d, a), o4 = orient(c, d, b);
  if (o1 != o2 && o3 != o4 && o1 * o2 <= 0 && o3 * o4 <= 0) {
    if (o1 != 0 || o2 != 0) return true;
  }
  if (o1 == 0 && on_segment_collinear(a, b, c)) return true;
  if (o2 == 0 && on_segment_collinear(a, b, d)) return true;
  if (o3 == 0 && on_segment_collinear(c, d, a)) return true;
  if (o4 == 0 && on_segment_collinear(c, d, b)) return true;
  return false;
}

// a single crossing point interior to both segments (geo's line_intersection
// with is_proper: no endpoint on the other segment, not collinear)
bool seg_proper(const P &a, const P &b, const P &c, const P &d) {
  const int o1 = orient(a, b, c), o2 = orient(a, b, d), o3 = orient(c, d, a), o4 = orient(c, d, b);
  return o1 != 0 && o2 != 0 && o3 != 0 && o4 != 0 && o1 != o2 && o3 != o4;
}

double dist_pt_seg(const P &p, const P &a, const P &b, P *closest = nullptr) {
  const double dx = b.x - a.x, dy = b.y - a.y;
  const double len2 = dx * dx + dy * dy;
  P q = a;
  if (len2 > 0.0) {
    double t = ((p.x - a.x) * dx + (p.y - a.y) * dy) / len2;
    if (t <= 0.0) {
      q = a;
    } else if (t >= 1.0) {
      q = b;
    } else {
      q = P{a.x + t * dx, a.y + t * dy};
    }
  }
  if (closest) *closest = q;
  return std::hypot(p.x - q.x, p.y - q.y);
}

double dist_seg_seg(const P &a, const P &b, const P &c, const P &d) {
  if (seg_intersects(a, b, c, d)) return 0.0;
  return std::min(std::min(dist_pt_seg(a, c, d), dist_pt_seg(b, c, d)),
                  std::min(dist_pt_seg(c, a, b), dist_pt_seg(d, a, b)));
}

struct Box {
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  void add(const P &p) {
    x0 = std::min(x0, p.x);
    y0 = std::min(y0, p.y);
    x1 = std::max(x1, p.x);
    y1 = std::max(y1, p.y);
  }
  bool hits(const Box &o) const { return x0 <= o.x1 && o.x0 <= x1 && y0 <= o.y1 && o.y0 <= y1; }
  double dist(const P &p) const {
    const double dx = std::max(std::max(x0 - p.x, 0.0), p.x - x1);
    const double dy = std::max(std::max(y0 - p.y, 0.0), p.y - y1);
    return std::hypot(dx, dy);
  }
};

// ---- geometries: parts (point, line, polygon with rings) ---------------------
enum Kind : uint8_t { kPoint = 0, kLine = 1, kPolygon = 2 };

struct Part {
  Kind kind;
  std::vector<std::vector<P>> rings;  // one ring for points and lines
};

struct Geom {
  std::vector<Part> parts;
  Box box;
  bool is_point() const { return parts.size() == 1 && parts[0].kind == kPoint; }
};

std::vector<Geom> read_geoms(const cs_geoms *g) {
  std::vector<Geom> out;
  if (!g) return out;
  out.resize(g->n_geoms);
  for (uint32_t i = 0; i < g->n_geoms; ++i) {
    for (uint32_t p = g->geom_part_off[i]; p < g->geom_part_off[i + 1]; ++p) {
      Part part;
      if (g->part_kind[p] > kPolygon) cs::fail(CS_EINVAL, "unknown geometry part kind");
      part.kind = (Kind)g->part_kind[p];
      for (uint32_t r = g->part_ring_off[p]; r < g->part_ring_off[p + 1]; ++r) {
        std::vector<P> ring;
        for (uint32_t k = g->ring_off[r]; k < g->ring_off[r + 1]; ++k) {
          const P q{g->xy[2 * (size_t)k], g->xy[2 * (size_t)k + 1]};
          if (!std::isfinite(q.x) || !std::isfinite(q.y)) cs::fail(CS_EINVAL, "geometry coordinates must be finite");
          ring.push_back(q);
          out[i].box.add(q);
        }
        // polygon rings close on their first point
        if (part.kind == kPolygon && ring.size() > 1 && (ring.front().x != ring.back().x || ring.front().y != ring.back().y))
          ring.push_back(ring.front());
        part.rings.push_back(std::move(ring));
      }
      if (part.rings.empty() || part.rings[0].empty()) continue;  // empty parts carry nothing
      out[i].parts.push_back(std::move(part));
    }
  }
  return out;
}

template <typename F>
void for_segments(const std::vector<P> &r, F f) {
  for (size_t k = 0; k + 1 < r.size(); ++k) f(r[k], r[k + 1]);
}

// point inside a polygon (boundary included), holes excluded
bool in_polygon(const Part &poly, const P &p) {
  for (const auto &ring : poly.rings)
    for (size_t k = 0; k + 1 < ring.size(); ++k)
      if (orient(ring[k], ring[k + 1], p) == 0 && on_segment_collinear(ring[k], ring[k + 1], p)) return true;
  auto inside_ring = [&](const std::vector<P> &ring) {
    bool in = false;
    for (size_t k = 0, j = ring.size() - 1; k < ring.size(); j = k++) {
      const P &a = ring[k], &b = ring[j];
      if ((a.y > p.y) != (b.y > p.y)) {
        const int o = orient(a, b, p);
        // the edge crosses the horizontal line through p to the right of p: p
        // lies left of the edge taken upwards
        if ((b.y > a.y) ? o > 0 : o < 0) in = !in;
      }
    }
    return in;
  };
  if (!inside_ring(poly.rings[0])) return false;
  for (size_t h = 1; h < poly.rings.size(); ++h)
    if (inside_ring(poly.rings[h])) return false;
  return true;
}

// Euclidean distance from a point to a geometry (0 inside a polygon), and the
// nearest point of the geometry
double dist_point_geom(const P &p, const Geom &g, P *closest) {
  double best = INFINITY;
  P bq{0.0, 0.0};
  for (const Part &part : g.parts) {
    if (part.kind == kPolygon && in_polygon(part, p)) {
      if (closest) *closest = p;
      return 0.0;
    }
    for (const auto &ring : part.rings) {
      if (part.kind == kPoint || ring.size() == 1) {
        const double d = std::hypot(p.x - ring[0].x, p.y - ring[0].y);
        if (d < best) {
          best = d;
          bq = ring[0];
        }
        continue;
      }
      for_segments(ring, [&](const P &a, const P &b) {
        P q;
        const double d = dist_pt_seg(p, a, b, &q);
        if (d < best) {
          best = d;
          bq = q;
        }
      });
    }
  }
  if (closest) *closest = bq;
  return best;
}

// Euclidean distance from a geometry to a street polyline (0 when they meet)
double dist_geom_line(const Geom &g, const std::vector<P> &line) {
  double best = INFINITY;
  for (const Part &part : g.parts) {
    if (part.kind == kPolygon)
      for (const P &q : line)
        if (in_polygon(part, q)) return 0.0;
    for (const auto &ring : part.rings) {
      if (part.kind == kPoint || ring.size() == 1) {
        for_segments(line, [&](const P &a, const P &b) { best = std::min(best, dist_pt_seg(ring[0], a, b)); });
        if (line.size() == 1) best = std::min(best, std::hypot(ring[0].x - line[0].x, ring[0].y - line[0].y));
        continue;
      }
      for_segments(ring, [&](const P &a, const P &b) {
        for_segments(line, [&](const P &c, const P &d) { best = std::min(best, dist_seg_seg(a, b, c, d)); });
      });
    }
    if (best == 0.0) return 0.0;
  }
  return best;
}

// the segment [a, b] meets a barrier geometry (line.intersects(geom))
bool seg_meets_geom(const P &a, const P &b, const Geom &g) {
  for (const Part &part : g.parts) {
    if (part.kind == kPolygon && (in_polygon(part, a) || in_polygon(part, b))) return true;
    for (const auto &ring : part.rings) {
      if (part.kind == kPoint || ring.size() == 1) {
        if (orient(a, b, ring[0]) == 0 && on_segment_collinear(a, b, ring[0])) return true;
        continue;
      }
      for (size_t k = 0; k + 1 < ring.size(); ++k)
        if (seg_intersects(a, b, ring[k], ring[k + 1])) return true;
    }
  }
  return false;
}

// ---- street edges in a uniform grid of envelopes ------------------------------
struct Street {
  uint32_t start, end;
  std::vector<P> line;
  Box box;
};

struct Grid {
  double x0 = 0, y0 = 0, cell = 1;
  int64_t nx = 1, ny = 1;
  std::vector<std::vector<uint32_t>> cells;
  int64_t cx(double x) const { return std::clamp<int64_t>((int64_t)std::floor((x - x0) / cell), 0, nx - 1); }
  int64_t cy(double y) const { return std::clamp<int64_t>((int64_t)std::floor((y - y0) / cell), 0, ny - 1); }
  void build(const std::vector<Street> &s) {
    Box all;
    double span = 0.0;
    for (const Street &e : s) {
      all.add(P{e.box.x0, e.box.y0});
      all.add(P{e.box.x1, e.box.y1});
      span += std::max(e.box.x1 - e.box.x0, e.box.y1 - e.box.y0);
    }
    x0 = all.x0;
    y0 = all.y0;
    const double w = std::max(all.x1 - all.x0, 1e-9), h = std::max(all.y1 - all.y0, 1e-9);
    // about as many cells as edges, none smaller than the mean edge envelope
    cell = std::max({span / std::max<size_t>(s.size(), 1), std::sqrt(w * h / std::max<size_t>(s.size(), 1)), 1e-6});
    nx = std::min<int64_t>(std::max<int64_t>((int64_t)(w / cell) + 1, 1), 1 << 14);
    ny = std::min<int64_t>(std::max<int64_t>((int64_t)(h / cell) + 1, 1), 1 << 14);
    cell = std::max(cell, std::max(w / nx, h / ny) * (1.0 + 1e-12));
    cells.assign((size_t)(nx * ny), {});
    for (uint32_t i = 0; i < s.size(); ++i)
      for (int64_t y = cy(s[i].box.y0); y <= cy(s[i].box.y1); ++y)
        for (int64_t x = cx(s[i].box.x0); x <= cx(s[i].box.x1); ++x) cells[(size_t)(y * nx + x)].push_back(i);
  }
  // edges whose envelope meets the box, each once, ascending
  void query(const Box &b, std::vector<uint32_t> &out, std::vector<uint8_t> &seen,
             const std::vector<Street> &s) const {
    out.clear();
    for (int64_t y = cy(b.y0); y <= cy(b.y1); ++y)
      for (int64_t x = cx(b.x0); x <= cx(b.x1); ++x)
        for (uint32_t i : cells[(size_t)(y * nx + x)])
          if (!seen[i] && s[i].box.hits(b)) {
            seen[i] = 1;
            out.push_back(i);
          }
    for (uint32_t i : out) seen[i] = 0;
    std::sort(out.begin(), out.end());
  }
  // the n edges of smallest envelope distance (ties: lower edge first)
  void nearest(const P &p, uint32_t n, std::vector<std::pair<double, uint32_t>> &out, std::vector<uint8_t> &seen,
               const std::vector<Street> &s) const {
    out.clear();
    if (s.empty() || n == 0) return;
    std::vector<uint32_t> touched;
    const int64_t px = cx(p.x), py = cy(p.y);
    const int64_t rmax = std::max(nx, ny);
    for (int64_t r = 0; r <= rmax; ++r) {
      for (int64_t y = py - r; y <= py + r; ++y) {
        if (y < 0 || y >= ny) continue;
        for (int64_t x = px - r; x <= px + r; ++x) {
          if (x < 0 || x >= nx) continue;
          if (std::max(std::llabs(x - px), std::llabs(y - py)) != r) continue;  // ring r only
          for (uint32_t i : cells[(size_t)(y * nx + x)]) {
            if (seen[i]) continue;
            seen[i] = 1;
            touched.push_back(i);
            out.emplace_back(s[i].box.dist(p), i);
          }
        }
      }
      // every edge not yet met has its envelope outside the searched square:
      // farther than the square's inner margin from p
      const double margin = std::min({p.x - (x0 + (px - r) * cell), (x0 + (px + r + 1) * cell) - p.x,
                                      p.y - (y0 + (py - r) * cell), (y0 + (py + r + 1) * cell) - p.y});
      if (out.size() >= n) {
        std::nth_element(out.begin(), out.begin() + (n - 1), out.end());
        const double nth = out[n - 1].first;
        if (nth < margin || (px - r <= 0 && py - r <= 0 && px + r >= nx - 1 && py + r >= ny - 1)) break;
      } else if (px - r <= 0 && py - r <= 0 && px + r >= nx - 1 && py + r >= ny - 1) {
        break;
      }
    }
    for (uint32_t i : touched) seen[i] = 0;
    std::sort(out.begin(), out.end());
    if (out.size() > n) out.resize(n);
  }
};

struct Assign {
  uint32_t data, node;
  double dist;
};

}  // namespace

struct cs_assignment {
  std::vector<Assign> rows;
};

extern "C" int cs_assign_data(uint32_t n_nodes, const double *node_xy, uint32_t n_edges, const uint32_t *edge_start,
                              const uint32_t *edge_end, const uint32_t *edge_off, const double *edge_xy,
                              const cs_geoms *barriers, const cs_geoms *data, double max_dist,
                              uint32_t n_nearest, cs_assignment **out, uint64_t *n_out) {
  return cs::guarded_call([&] {
    if (!out || !n_out) cs::fail(CS_EINVAL, "out and n_out must not be NULL");
    *out = nullptr;
    *n_out = 0;
    if (!data) cs::fail(CS_EINVAL, "data geometries must not be NULL");
    if (n_nodes && !node_xy) cs::fail(CS_EINVAL, "node coordinates must not be NULL");
    if (n_edges && (!edge_start || !edge_end || !edge_off || !edge_xy))
      cs::fail(CS_EINVAL, "edge arrays must not be NULL");
    if (!(max_dist >= 0.0)) cs::fail(CS_EINVAL, "max_assignment_dist must be non-negative");
    // build_edge_rtree (graph.rs:1064-1146): one entry per node pair and
    // geometry, either direction, in edge order
    std::vector<Street> streets;
    std::unordered_map<uint64_t, std::vector<uint32_t>> seen_pair;
    for (uint32_t e = 0; e < n_edges; ++e) {
      if (edge_start[e] >= n_nodes || edge_end[e] >= n_nodes) cs::fail(CS_EINVAL, "edge node out of range");
      Street s;
      s.start = edge_start[e];
      s.end = edge_end[e];
      for (uint32_t k = edge_off[e]; k < edge_off[e + 1]; ++k) {
        const P q{edge_xy[2 * (size_t)k], edge_xy[2 * (size_t)k + 1]};
        s.line.push_back(q);
        s.box.add(q);
      }
      if (s.line.size() < 2) cs::fail(CS_EINVAL, "street edge geometry needs at least 2 coordinates");
      const uint64_t key = ((uint64_t)std::min(s.start, s.end) << 32) | std::max(s.start, s.end);
      auto &prev = seen_pair[key];
      bool dup = false;
      for (uint32_t o : prev) {
        const auto &l = streets[o].line;
        if (l.size() != s.line.size()) continue;
        bool fw = true, bw = true;
        for (size_t k = 0; k < l.size(); ++k) {
          fw = fw && l[k].x == s.line[k].x && l[k].y == s.line[k].y;
          const P &r = s.line[s.line.size() - 1 - k];
          bw = bw && l[k].x == r.x && l[k].y == r.y;
        }
        if (fw || bw) dup = true;
      }
      if (dup) continue;
      prev.push_back((uint32_t)streets.size());
      streets.push_back(std::move(s));
    }
    std::vector<Geom> bars = read_geoms(barriers);
    std::vector<Geom> items = read_geoms(data);
    Grid grid;
    grid.build(streets);
    auto node_pt = [&](uint32_t v) { return P{node_xy[2 * (size_t)v], node_xy[2 * (size_t)v + 1]}; };

    // per-thread scratch of the envelope queries
    struct Scratch {
      std::vector<uint8_t> seen;
      std::vector<std::pair<double, uint32_t>> near;
    };

    // the line from the data's nearest point to a node crosses a barrier or a
    // street properly (graph.rs:1488-1535)
    auto blocked = [&](const P &a, const P &b, Scratch &sc) {
      Box lb;
      lb.add(a);
      lb.add(b);
      for (const Geom &g : bars)
        if (g.box.hits(lb) && seg_meets_geom(a, b, g)) return true;
      if (streets.empty()) return false;
      std::vector<uint32_t> cand;
      grid.query(lb, cand, sc.seen, streets);
      for (uint32_t i : cand) {
        const auto &l = streets[i].line;
        for (size_t k = 0; k + 1 < l.size(); ++k)
          if (seg_proper(a, b, l[k], l[k + 1])) return true;
      }
      return false;
    };

    // the assignments of data geometry d, appended to rows in the reference's order
    auto assign_one = [&](uint32_t d, Scratch &sc, std::vector<Assign> &rows) {
      const Geom &g = items[d];
      if (g.parts.empty() || streets.empty()) return;  // no geometry, no network: nothing assigned
      // candidates (graph.rs:1367-1411)
      std::vector<uint32_t> cands;
      if (g.is_point()) {
        grid.nearest(g.parts[0].rings[0][0], n_nearest, sc.near, sc.seen, streets);
        for (auto &q : sc.near) cands.push_back(q.second);
      } else {
        Box qb = g.box;
        qb.x0 -= max_dist;
        qb.y0 -= max_dist;
        qb.x1 += max_dist;
        qb.y1 += max_dist;
        grid.query(qb, cands, sc.seen, streets);
      }
      std::vector<std::pair<double, uint32_t>> within;
      for (uint32_t i : cands) {
        const double dd = dist_geom_line(g, streets[i].line);
        if (dd <= max_dist) within.emplace_back(dd, i);
      }
      std::stable_sort(within.begin(), within.end(),
                       [](const auto &x, const auto &y) { return x.first < y.first; });
      // node checks, cached per entry (graph.rs:1413-1486)
      std::unordered_map<uint32_t, std::pair<bool, double>> checked;
      std::unordered_set<uint32_t> added;
      auto check = [&](uint32_t v) -> std::pair<bool, double> {
        auto it = checked.find(v);
        if (it != checked.end()) return it->second;
        const P np = node_pt(v);
        P cp;
        const double nd = dist_point_geom(np, g, &cp);
        std::pair<bool, double> r;
        if (nd < 1e-6) r = {true, 0.0};
        else r = {!blocked(cp, np, sc), nd};
        checked.emplace(v, r);
        return r;
      };
      uint32_t visited = 0;
      for (auto &w : within) {
        const Street &s = streets[w.second];
        bool produced = false;
        for (uint32_t v : {s.start, s.end}) {
          const auto r = check(v);
          if (r.first && added.insert(v).second) {
            rows.push_back(Assign{d, v, r.second});
            produced = true;
          }
        }
        if (g.is_point() && produced) break;
        if (++visited >= n_nearest) break;
      }
    };

    // entries are independent (the reference's rayon loop, data.rs:469-484):
    // contiguous blocks of them go to host threads, whose rows concatenate in
    // data order
    const uint32_t n_items = (uint32_t)items.size();
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t n_threads = std::max<uint32_t>(1u, std::min<uint32_t>({hw, 16u, n_items / 64}));
    std::vector<std::vector<Assign>> rows(n_threads);
    std::vector<std::exception_ptr> errs(n_threads);
    auto run_block = [&](uint32_t t) {
      try {
        Scratch sc;
        sc.seen.assign(streets.size(), 0);
        const uint32_t d0 = (uint32_t)((uint64_t)n_items * t / n_threads);
        const uint32_t d1 = (uint32_t)((uint64_t)n_items * (t + 1) / n_threads);
        for (uint32_t d = d0; d < d1; ++d) assign_one(d, sc, rows[t]);
      } catch (...) {
        errs[t] = std::current_exception();
      }
    };
    {
      std::vector<std::thread> pool;
      for (uint32_t t = 1; t < n_threads; ++t) pool.emplace_back(run_block, t);
      run_block(0);
      for (auto &th : pool) th.join();
    }
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);
    auto res = std::make_unique<cs_assignment>();
    for (auto &r : rows) res->rows.insert(res->rows.end(), r.begin(), r.end());
    *n_out = res->rows.size();
    *out = res.release();
  });
}

extern "C" int cs_assignment_copy(const cs_assignment *a, uint32_t *data_index, uint32_t *node, double *dist) {
  return cs::guarded_call([&] {
    if (!a) cs::fail(CS_EINVAL, "assignment must not be NULL");
    for (size_t i = 0; i < a->rows.size(); ++i) {
      if (data_index) data_index[i] = a->rows[i].data;
      if (node) node[i] = a->rows[i].node;
      if (dist) dist[i] = a->rows[i].dist;
    }
  });
}

extern "C" void cs_assignment_free(cs_assignment *a) { delete a; }
