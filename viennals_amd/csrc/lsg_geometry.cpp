// Initial level sets (the MakeGeometry part of the reference) built on the
// host as lists of defined points; lsg_host turns them into device bands.
#include "lsg_geometry.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace lsg {

// MakeGeometry::makeSphere (lsMakeGeometry.hpp:177-241): along-axis distance
// to the sphere, kept where |distance| <= gridDelta + 1e-5, visited x fastest.
// Bit d of ignoreBcMask: setIgnoreBoundaryConditions for dimension d.
void sphereBand(int D, const GridInfo &g, const double *origin, double radius, std::vector<int32_t> &idx,
                std::vector<double> &val, int ignoreBcMask) {
  const double gridDelta = g.dx;
  int32_t index[3] = {0, 0, 0}, endIndex[3] = {0, 0, 0};
  for (int i = 0; i < D; ++i) {
    index[i] = (int32_t)((origin[i] - radius) / gridDelta - 1);
    endIndex[i] = (int32_t)((origin[i] + radius) / gridDelta + 1);
  }
  const double initialWidth = 2.;
  const double valueLimit = initialWidth * 0.5 * gridDelta + 1e-5;
  const double radius2 = radius * radius;
  int32_t minIndex[3] = {index[0], index[1], index[2]};
  auto less = [D](const int32_t *a, const int32_t *b) {
    for (int d = D - 1; d >= 0; --d)
      if (a[d] != b[d]) return a[d] < b[d];
    return false;
  };
  while (less(index, endIndex)) {
    double distance = std::numeric_limits<double>::max();
    for (int i = 0; i < D; ++i) {
      double y = (index[(i + 1) % D] * gridDelta) - origin[(i + 1) % D];
      double z = 0;
      if (D == 3) z = (index[(i + 2) % D] * gridDelta) - origin[(i + 2) % D];
      double x = radius2 - y * y - z * z;
      if (x < 0.) continue;
      double dirRadius = std::fabs((index[i] * gridDelta) - origin[i]) - std::sqrt(x);
      if (std::fabs(dirRadius) < std::fabs(distance)) distance = dirRadius;
    }
    if (std::fabs(distance) <= valueLimit) {
      for (int d = 0; d < D; ++d) {
        int32_t c = index[d];
        // periodic indices are mapped into the grid unless the boundary
        // condition is ignored (lsMakeGeometry.hpp:227-235)
        if (g.bc[d] == PERIODIC && !((ignoreBcMask >> d) & 1)) c = g.map(d, c);
        idx.push_back(c);
      }
      val.push_back(distance / gridDelta);
    }
    int dim = 0;
    for (; dim < D - 1; ++dim) {
      if (index[dim] < endIndex[dim]) break;
      index[dim] = minIndex[dim];
    }
    ++index[dim];
  }
}

// Plane through origin with the given normal (lsMakeGeometry.hpp:245-360):
// signed Euclidean distance in grid units, kept where |phi| <= 1.  The plane
// must be transversal to the single infinite dimension, which bounds the
// band; the other dimensions span the grid.
int planeBand(int D, const GridInfo &g, const double *origin, const double *normal, std::vector<int32_t> &idx,
              std::vector<double> &val) {
  double n[3] = {0, 0, 0}, mod = 0;
  for (int i = 0; i < D; ++i) mod += normal[i] * normal[i];
  mod = std::sqrt(mod);
  if (!(mod > 0)) return -1;
  for (int i = 0; i < D; ++i) n[i] = normal[i] / mod;
  int inf = -1;
  for (int d = 0; d < D; ++d)
    if (!g.bounded(d)) {
      if (inf >= 0) return -2;
      inf = d;
    }
  if (inf < 0 || n[inf] == 0.) return -3;
  const double dx = g.dx;
  // loop over the bounded dims, solve for the range of the infinite dim
  int other[2] = {-1, -1}, no = 0;
  for (int d = 0; d < D; ++d)
    if (d != inf) other[no++] = d;
  struct P {
    int32_t c[3];
    double v;
  };
  std::vector<P> pts;
  int32_t lo1 = g.lo[other[0]], hi1 = g.hi[other[0]];
  int32_t lo2 = (D == 3) ? g.lo[other[1]] : 0, hi2 = (D == 3) ? g.hi[other[1]] : 0;
  for (int32_t b = lo2; b <= hi2; ++b)
    for (int32_t a = lo1; a <= hi1; ++a) {
      double s = (a * dx - origin[other[0]]) * n[other[0]];
      if (D == 3) s += (b * dx - origin[other[1]]) * n[other[1]];
      // phi(c) = (s + (c*dx - o)*n_inf) / dx ; |phi| <= 1
      double c0 = origin[inf] + (-dx - s) / n[inf];
      double c1 = origin[inf] + (dx - s) / n[inf];
      if (c0 > c1) std::swap(c0, c1);
      int32_t k0 = (int32_t)std::floor(c0 / dx) - 1, k1 = (int32_t)std::ceil(c1 / dx) + 1;
      for (int32_t k = k0; k <= k1; ++k) {
        double phi = (s + (k * dx - origin[inf]) * n[inf]) / dx;
        if (std::fabs(phi) <= 1.0) {
          P p;
          p.c[other[0]] = a;
          if (D == 3) p.c[other[1]] = b;
          p.c[inf] = k;
          if (D == 2) p.c[2] = 0;
          p.v = phi;
          pts.push_back(p);
        }
      }
    }
  std::sort(pts.begin(), pts.end(), [D](const P &x, const P &y) {
    for (int d = D - 1; d >= 0; --d)
      if (x.c[d] != y.c[d]) return x.c[d] < y.c[d];
    return false;
  });
  for (auto &p : pts) {
    for (int d = 0; d < D; ++d) idx.push_back(p.c[d]);
    val.push_back(p.v);
  }
  return 0;
}

// Axis-aligned box (lsMakeGeometry.hpp:363-419): signed Euclidean distance
// to the box surface in grid units, kept where |phi| <= 1.
void boxBand(int D, const GridInfo &g, const double *mn, const double *mx, std::vector<int32_t> &idx,
             std::vector<double> &val) {
  const double dx = g.dx;
  int32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < D; ++d) {
    lo[d] = (int32_t)std::floor(mn[d] / dx) - 2;
    hi[d] = (int32_t)std::ceil(mx[d] / dx) + 2;
    if (g.bounded(d)) {
      lo[d] = std::max(lo[d], g.lo[d]);
      hi[d] = std::min(hi[d], g.hi[d]);
    }
  }
  int32_t c[3];
  for (c[2] = lo[2]; c[2] <= hi[2]; ++c[2])
    for (c[1] = lo[1]; c[1] <= hi[1]; ++c[1])
      for (c[0] = lo[0]; c[0] <= hi[0]; ++c[0]) {
        double out2 = 0, inside = std::numeric_limits<double>::max();
        bool in = true;
        for (int d = 0; d < D; ++d) {
          double p = c[d] * dx;
          double e = 0;
          if (p < mn[d]) e = mn[d] - p, in = false;
          else if (p > mx[d]) e = p - mx[d], in = false;
          out2 += e * e;
          double di = std::min(p - mn[d], mx[d] - p);
          inside = std::min(inside, di);
        }
        double phi = in ? -inside / dx : std::sqrt(out2) / dx;
        if (std::fabs(phi) <= 1.0) {
          for (int d = 0; d < D; ++d) idx.push_back(c[d]);
          val.push_back(phi);
        }
      }
}

}  // namespace lsg
