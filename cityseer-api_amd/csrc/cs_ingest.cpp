// cs_ingest.cpp -- street-network ingest on the host: the edge payload that
// NetworkStructure.add_street_edge derives from an edge's polyline
// (rust/src/graph.rs:728-889), for every edge of a network in one call, and
// the splitting of polylines into pieces no longer than a limit, as
// tools.graphs.nx_decompose does (pysrc/cityseer/tools/graphs.py:1829-1962).
// Both run before any device work and need no GPU.

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "cs_api.hpp"

namespace {

struct XY {
  double x, y;
};

inline bool same(const XY &a, const XY &b) { return a.x == b.x && a.y == b.y; }

// measure_bearing (graph.rs:326-332): degrees, 0 for coincident points
inline double bearing(const XY &a, const XY &b) {
  if (same(a, b)) return 0.0;
  return std::atan2(b.y - a.y, b.x - a.x) * (180.0 / M_PI);
}

// measure_coords_angle (graph.rs:336-346)
inline double turn(const XY &a, const XY &b, const XY &c) {
  if (same(a, b) || same(b, c)) return 0.0;
  const double d = bearing(c, b) - bearing(b, a);
  double r = std::fmod(d + 180.0, 360.0);  // rem_euclid
  if (r < 0.0) r += 360.0;
  return std::fabs(r - 180.0);
}

// Decomposed edges: each edge's pieces' polylines back to back, `dims`
// doubles per point (x, y and, for 3-D input, z).
struct Pieces {
  uint32_t dims = 2;
  std::vector<uint32_t> edge_cuts;  // pieces per input edge
  std::vector<uint32_t> piece_off;  // [n_pieces + 1] into the points
  std::vector<double> pts;          // [n_points][dims]
};

// Length of segment i of a polyline with `dims` doubles per point, as shapely
// (GEOS) measures it when splitting: the square root of the summed squares
// of the planar differences.
inline double seg_len(const double *c, uint32_t dims, uint32_t i) {
  const double dx = c[(i + 1) * dims] - c[i * dims], dy = c[(i + 1) * dims + 1] - c[i * dims + 1];
  return std::sqrt(dx * dx + dy * dy);
}

// The point at distance `dist` along the polyline, appended to `out`, as GEOS
// interpolates (LengthIndexedLine): the first segment that extends strictly
// beyond `dist`, at fraction (dist - walked) / length of it, every coordinate
// (z too) interpolated alike; the last point when none does. A distance that
// falls exactly on a vertex therefore yields that vertex itself.
void point_at(const double *c, uint32_t n, uint32_t dims, const std::vector<double> &sl, double dist,
              std::vector<double> &out) {
  double walked = 0.0;
  for (uint32_t i = 0; i + 1 < n; ++i) {
    if (walked + sl[i] > dist) {
      const double t = (dist - walked) / sl[i];
      for (uint32_t d = 0; d < dims; ++d) out.push_back(c[i * dims + d] + t * (c[(i + 1) * dims + d] - c[i * dims + d]));
      return;
    }
    walked += sl[i];
  }
  for (uint32_t d = 0; d < dims; ++d) out.push_back(c[(n - 1) * dims + d]);
}

// shapely ops.substring(line, start, end) for start < end: the interpolated
// start, the vertices whose distance along the line lies strictly between,
// the interpolated end.
void substring(const double *c, uint32_t n, uint32_t dims, const std::vector<double> &sl, double total,
               double start, double end, std::vector<double> &out) {
  start = std::min(std::max(start, 0.0), total);
  end = std::min(std::max(end, 0.0), total);
  point_at(c, n, dims, sl, start, out);
  double walked = 0.0;
  for (uint32_t i = 0; i + 1 < n; ++i) {
    walked += sl[i];
    if (walked >= end) break;
    if (start < walked)
      for (uint32_t d = 0; d < dims; ++d) out.push_back(c[(i + 1) * dims + d]);
  }
  point_at(c, n, dims, sl, end, out);
}

}  // namespace

struct cs_pieces {
  Pieces p;
};

// Cuts every polyline into equal pieces: ceil(length / decompose_max) of them,
// or exactly two (at half its length) when halve is set.
static void cut_lines(uint32_t n_edges, const uint32_t *coord_off, const double *pts, uint32_t dims,
                      double decompose_max, bool halve, cs_pieces **out, uint32_t *n_pieces, uint32_t *n_points) {
  if (!out || !n_pieces || !n_points) cs::fail(CS_EINVAL, "out, n_pieces and n_points must not be NULL");
  if (n_edges && (!coord_off || !pts)) cs::fail(CS_EINVAL, "coord_off and pts must not be NULL");
  if (dims != 2 && dims != 3) cs::fail(CS_EINVAL, "dims must be 2 or 3");
  if (!halve && (!(decompose_max > 0.0) || !std::isfinite(decompose_max)))
    cs::fail(CS_EINVAL, "decompose_max must be finite and positive");
  {
    auto h = std::make_unique<cs_pieces>();
    Pieces &P = h->p;
    P.dims = dims;
    P.edge_cuts.resize(n_edges);
    P.piece_off.push_back(0);
    std::vector<double> sl;
    for (uint32_t e = 0; e < n_edges; ++e) {
      const uint32_t b = coord_off[e], n = coord_off[e + 1] - coord_off[e];
      if (coord_off[e + 1] < b || n < 2)
        cs::fail(CS_EINVAL, "edge " + std::to_string(e) + " needs at least 2 coordinates");
      const double *c = pts + (size_t)b * dims;
      sl.assign(n - 1, 0.0);
      double total = 0.0;
      for (uint32_t i = 0; i + 1 < n; ++i) {
        sl[i] = seg_len(c, dims, i);
        total += sl[i];
      }
      if (!(total > 0.0)) cs::fail(CS_EINVAL, "edge " + std::to_string(e) + " has zero length");
      // cuts = ceil(length / decompose_max) equal steps; the last piece runs to
      // the end of the line (graphs.py:1911-1961). Halves: [0, total/2], [total/2, total]
      const uint32_t cuts = halve ? 2u : (uint32_t)std::max(1.0, std::ceil(total / decompose_max));
      const double step = halve ? 0.5 * total : total / cuts;
      P.edge_cuts[e] = cuts;
      double at = 0.0;
      for (uint32_t k = 0; k < cuts; ++k) {
        const double end = (k + 1 == cuts) ? total : at + step;
        substring(c, n, dims, sl, total, at, end, P.pts);
        P.piece_off.push_back((uint32_t)(P.pts.size() / dims));
        at += step;
      }
    }
    *n_pieces = (uint32_t)(P.piece_off.size() - 1);
    *n_points = (uint32_t)(P.pts.size() / dims);
    *out = h.release();
  }
}

extern "C" {

int cs_edge_geometry(uint32_t n_edges, const uint32_t *coord_off, const double *xy, float *length, float *angle_sum,
                     float *in_bearing, float *out_bearing) {
  return cs::guarded_call([&] {
    if (n_edges && (!coord_off || !xy)) cs::fail(CS_EINVAL, "coord_off and xy must not be NULL");
    for (uint32_t e = 0; e < n_edges; ++e) {
      const uint32_t b = coord_off[e], n = coord_off[e + 1] - coord_off[e];
      if (coord_off[e + 1] < b || n < 2)
        cs::fail(CS_EINVAL, "Street edge geometry (idx " + std::to_string(e) +
                                ") must have at least 2 coordinates. Found " + std::to_string(n < 2 ? n : 0) + ".");
      const XY *c = reinterpret_cast<const XY *>(xy) + b;
      // Euclidean length of the LineString (graph.rs:774): hypot per segment, summed in f64
      double len = 0.0;
      for (uint32_t i = 0; i + 1 < n; ++i) len += std::hypot(c[i + 1].x - c[i].x, c[i + 1].y - c[i].y);
      if (length) length[e] = (float)len;
      // cumulative turn angle (graph.rs:368-379, 857)
      double ang = 0.0;
      for (uint32_t i = 1; i + 1 < n; ++i) ang += turn(c[i - 1], c[i], c[i + 1]);
      if (angle_sum) angle_sum[e] = (float)ang;
      // first / last distinct coordinate pairs (graph.rs:797-855); 0 for degenerate edges
      if (in_bearing) {
        float v = 0.0f;
        for (uint32_t i = 1; i < n; ++i)
          if (!same(c[i], c[0])) {
            v = (float)bearing(c[0], c[i]);
            break;
          }
        in_bearing[e] = v;
      }
      if (out_bearing) {
        float v = 0.0f;
        for (uint32_t i = n - 1; i-- > 0;)
          if (!same(c[i], c[n - 1])) {
            v = (float)bearing(c[i], c[n - 1]);
            break;
          }
        out_bearing[e] = v;
      }
    }
  });
}

int cs_decompose(uint32_t n_edges, const uint32_t *coord_off, const double *pts, uint32_t dims,
                 double decompose_max, cs_pieces **out, uint32_t *n_pieces, uint32_t *n_points) {
  return cs::guarded_call(
      [&] { cut_lines(n_edges, coord_off, pts, dims, decompose_max, false, out, n_pieces, n_points); });
}

int cs_halve(uint32_t n_edges, const uint32_t *coord_off, const double *pts, uint32_t dims, cs_pieces **out,
             uint32_t *n_pieces, uint32_t *n_points) {
  return cs::guarded_call([&] { cut_lines(n_edges, coord_off, pts, dims, 0.0, true, out, n_pieces, n_points); });
}

int cs_pieces_copy(const cs_pieces *h, uint32_t *edge_cuts, uint32_t *piece_off, double *pts) {
  return cs::guarded_call([&] {
    if (!h) cs::fail(CS_EINVAL, "pieces must not be NULL");
    const Pieces &P = h->p;
    if (edge_cuts) std::memcpy(edge_cuts, P.edge_cuts.data(), 4 * P.edge_cuts.size());
    if (piece_off) std::memcpy(piece_off, P.piece_off.data(), 4 * P.piece_off.size());
    if (pts) std::memcpy(pts, P.pts.data(), 8 * P.pts.size());
  });
}

void cs_pieces_free(cs_pieces *h) { delete h; }

}  // extern "C"
