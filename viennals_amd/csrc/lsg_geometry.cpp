// Initial level sets (the MakeGeometry part of the reference) built on the
// host as lists of defined points; lsg_host turns them into device bands.
#include "lsg_geometry.hpp"

#include <algorithm>
#include <array>
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

// ------------------------------------------------- FromSurfaceMesh ----
//
// MakeGeometry builds planes and boxes as surface meshes and converts them
// with FromSurfaceMesh (lsMakeGeometry.hpp:245-419, lsFromSurfaceMesh.hpp:
// 243-480): every grid line parallel to an axis that crosses a mesh element
// gives the grid points next to the crossing their distance to it along that
// line (times 1 - 1e-7, clipped to +-1), and each point keeps the value of
// the crossing with the smallest sign distance.  The grid functions of
// ViennaHRLE (absent from the reference tree) are taken as: grid position =
// index * gridDelta, local index = coordinate / gridDelta, min / max index =
// the grid's end points, +-715827882 along infinite axes (the grid-minimum
// run start of the reference's BooleanOperationExactZero fixture) and the
// period's end (max grid point + 1) along periodic ones, and indices wrap
// along periodic axes.

namespace {

using Node = std::array<double, 3>;

struct MeshPoint {
  std::array<int32_t, 3> c;
  double sign, real;
};

// lexicographic comparison of two nodes (VectorType<T, 3>::operator<)
bool nodeLess(const Node &a, const Node &b) {
  for (int i = 0; i < 3; ++i)
    if (a[i] != b[i]) return a[i] < b[i];
  return false;
}

// calculateGridlineTriangleIntersection for a line element in 2D
// (lsFromSurfaceMesh.hpp:118-155)
int crossing2(const Node &P, const Node *c, int dir, double &x) {
  bool inPos = true, inNeg = true;
  double A[2];
  const int dirA = (dir + 1) % 2;
  for (int k = 0; k < 2; ++k) {
    A[k] = c[(k + 1) % 2][dirA] - P[dirA];
    if (k == dir) A[k] = -A[k];
    if (A[k] < 0.) inPos = false;
    if (A[k] > 0.) inNeg = false;
  }
  if (!((!inPos && inNeg) || (inPos && !inNeg))) return 0;
  const double sum = A[0] + A[1];
  int k = 0;
  if (inPos ? (A[1] > A[0]) : (A[1] < A[0])) k = 1;
  x = c[k][dir] + (c[(k + 1) % 2][dir] - c[k][dir]) * (A[(k + 1) % 2] / sum);
  return inPos ? 1 : -1;
}

// calculateGridlineTriangleIntersection for a triangle in 3D
// (lsFromSurfaceMesh.hpp:163-215): signed areas of the sub-triangles seen
// along the line, each pair of nodes taken in lexicographic order so that
// neighbouring triangles agree on their shared edge
int crossing3(const Node &P, const Node *c, int dir, double &x) {
  bool inPos = true, inNeg = true;
  double A[3];
  const int dA = (dir + 1) % 3, dB = (dir + 2) % 3;
  for (int k = 0; k < 3; ++k) {
    const bool swapped = nodeLess(c[(k + 1) % 3], c[(k + 2) % 3]);
    const Node &v1 = swapped ? c[(k + 2) % 3] : c[(k + 1) % 3];
    const Node &v2 = swapped ? c[(k + 1) % 3] : c[(k + 2) % 3];
    A[k] = (v1[dA] - P[dA]) * (v2[dB] - P[dB]) - (v2[dA] - P[dA]) * (v1[dB] - P[dB]);
    if (swapped) A[k] = -A[k];
    if (A[k] < 0.) inPos = false;
    if (A[k] > 0.) inNeg = false;
  }
  if (!((!inPos && inNeg) || (inPos && !inNeg))) return 0;
  const double sum = A[0] + A[1] + A[2];
  int k = 0;
  for (int i = 1; i < 3; ++i)
    if (inPos ? (A[i] > A[k]) : (A[i] < A[k])) k = i;
  x = c[k][dir] + (c[(k + 1) % 3][dir] - c[k][dir]) * (A[(k + 1) % 3] / sum) +
      (c[(k + 2) % 3][dir] - c[k][dir]) * (A[(k + 2) % 3] / sum);
  return inPos ? 1 : -1;
}

// Grid::getMinIndex / getMaxIndex
int32_t minIndexOf(const GridInfo &g, int d) { return g.lo[d]; }
int32_t maxIndexOf(const GridInfo &g, int d) { return g.bc[d] == PERIODIC ? g.hi[d] + 1 : g.hi[d]; }

int32_t wrapPeriodic(const GridInfo &g, int d, int64_t i) {
  if (g.bc[d] != PERIODIC) return (int32_t)i;
  const int64_t L = (int64_t)g.hi[d] - g.lo[d] + 1;
  int64_t m = (i - g.lo[d]) % L;
  if (m < 0) m += L;
  return (int32_t)(g.lo[d] + m);
}

// FromSurfaceMesh::apply (lsFromSurfaceMesh.hpp:243-480) for elements of D
// nodes each; returns the defined points sorted in HRLE order.
void fromSurfaceMesh(int D, const GridInfo &g, const std::vector<Node> &nodesAll,
                     const std::vector<std::array<int, 3>> &elements, const bool *removeBoundaryTriangles,
                     std::vector<int32_t> &idx, std::vector<double> &val) {
  const double boundaryEps = 1e-5, distanceEps = 1e-4, signEps = 1e-6;
  const double dx = g.dx;
  bool removeB[3];
  for (int i = 0; i < D; ++i) removeB[i] = !(!removeBoundaryTriangles[i] && g.bc[i] == PERIODIC);
  auto maxGridPoint = [&](int i) { return g.hi[i]; };
  double gridMin[3], gridMax[3];
  for (int i = 0; i < D; ++i) {
    gridMin[i] = minIndexOf(g, i) * dx;
    gridMax[i] = maxIndexOf(g, i) * dx;
  }
  std::vector<MeshPoint> pts;
  for (const auto &el : elements) {
    Node nodes[3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double center[3] = {0, 0, 0};
    bool remove = false;
    for (int dim = 0; dim < D; ++dim) {
      bool atMin = true, atMax = true;  // flags[dim], flags[dim + D]
      for (int q = 0; q < D; ++q) {
        double v = nodesAll[el[q]][dim];
        if (std::fabs(v - gridMin[dim]) < boundaryEps * dx) v = gridMin[dim];
        if (std::fabs(v - gridMax[dim]) < boundaryEps * dx) v = gridMax[dim];
        nodes[q][dim] = v;
        if (v > gridMin[dim]) atMin = false;
        if (v < gridMax[dim]) atMax = false;
        center[dim] += v;
      }
      if (removeB[dim] && (atMin || atMax)) remove = true;
    }
    if (remove) continue;
    for (int q = 0; q < D; ++q) center[q] /= (double)D;
    Node minNode = nodes[0], maxNode = nodes[0];
    for (int i = 1; i < D; ++i)
      for (int q = 0; q < D; ++q) {
        minNode[q] = std::min(minNode[q], nodes[i][q]);
        maxNode[q] = std::max(maxNode[q], nodes[i][q]);
      }
    int32_t minIndex[3], maxIndex[3];
    for (int q = 0; q < D; ++q) {
      minIndex[q] = (int32_t)std::ceil(minNode[q] / dx);
      maxIndex[q] = (int32_t)std::floor(maxNode[q] / dx);
    }
    for (int i = 0; i < D; ++i)
      if (removeB[i]) {
        minIndex[i] = std::max(minIndex[i], minIndexOf(g, i));
        maxIndex[i] = std::min(maxIndex[i], maxGridPoint(i));
      }
    for (int z = 0; z < D; ++z) {
      // the grid lines along z through the element's index box; the box
      // orders its two corners per axis (lsFromSurfaceMesh.hpp:26-30), so an
      // element between two grid lines visits both of them
      const int a1 = (z + 1) % D, a2 = (z + 2) % D;
      const int32_t lo1 = std::min(minIndex[a1], maxIndex[a1]), hi1 = std::max(minIndex[a1], maxIndex[a1]);
      const int32_t lo2 = (D == 3) ? std::min(minIndex[a2], maxIndex[a2]) : 0;
      const int32_t hi2 = (D == 3) ? std::max(minIndex[a2], maxIndex[a2]) : 0;
      for (int32_t j2 = lo2; j2 <= hi2; ++j2)
        for (int32_t j1 = lo1; j1 <= hi1; ++j1) {
          int32_t it[3] = {0, 0, 0};
          it[a1] = j1;
          if (D == 3) it[a2] = j2;
          Node P = {0, 0, 0};
          P[a1] = it[a1] * dx;
          if (D == 3) P[a2] = it[a2] * dx;
          double x = 0;
          const int status = (D == 3) ? crossing3(P, nodes, z, x) : crossing2(P, nodes, z, x);
          if (status == 0) continue;
          if (removeB[z] && x > dx * maxIndexOf(g, z)) continue;
          if (removeB[z] && x < minIndexOf(g, z) * dx) continue;
          const double x2 = x / dx;
          int64_t lo = (int64_t)std::floor(x2 - distanceEps);
          int64_t hi = (int64_t)std::ceil(x2 + distanceEps);
          lo = std::max<int64_t>(lo, (int64_t)minIndex[z] - 1);
          hi = std::min<int64_t>(hi, (int64_t)maxIndex[z] + 1);
          lo = std::max<int64_t>(lo, minIndexOf(g, z));
          hi = std::min<int64_t>(hi, maxIndexOf(g, z));
          if (!removeB[z]) {
            lo = wrapPeriodic(g, z, lo);
            hi = wrapPeriodic(g, z, hi);
          }
          double t[3] = {0, 0, 0};
          for (int q = 0; q < D; ++q) t[q] = center[q];
          t[z] -= x;
          for (int k = 1; k < D; ++k) t[(z + k) % D] -= P[(z + k) % D];
          double tn = 0;
          for (int q = 0; q < D; ++q) tn += t[q] * t[q];
          tn = std::sqrt(tn);
          for (int q = 0; q < D; ++q) t[q] /= tn;
          for (int64_t k = lo; k <= hi; ++k) {
            double real = (double)k - x2;
            double sign = real;
            sign -= signEps * t[z];
            if (status < 0) {
              real = -real;
              sign = -sign;
            }
            if (real < 0.) sign = -sign;
            real *= (1. - distanceEps * 1e-3);
            if (real > 1.) real = 1.;
            if (real < -1.) real = -1.;
            if (real == 0.) real = 0.;
            MeshPoint m;
            m.c = {0, 0, 0};
            for (int q = 0; q < D; ++q) m.c[q] = wrapPeriodic(g, q, q == z ? k : it[q]);
            m.sign = sign;
            m.real = real;
            pts.push_back(m);
          }
        }
    }
  }
  // one value per point: the crossing with the smallest sign distance (then
  // the smallest distance), lsFromSurfaceMesh.hpp:457-470
  std::sort(pts.begin(), pts.end(), [D](const MeshPoint &a, const MeshPoint &b) {
    for (int d = D - 1; d >= 0; --d)
      if (a.c[d] != b.c[d]) return a.c[d] < b.c[d];
    if (a.sign < b.sign) return true;  // std::pair order of (sign, real)
    if (b.sign < a.sign) return false;
    return a.real < b.real;
  });
  for (size_t i = 0; i < pts.size(); ++i) {
    if (i > 0 && pts[i].c == pts[i - 1].c) continue;
    for (int d = 0; d < D; ++d) idx.push_back(pts[i].c[d]);
    val.push_back(pts[i].real);
  }
}

}  // namespace

// MakeGeometry::makePlane (lsMakeGeometry.hpp:245-360): the mesh of the
// plane over the grid, one grid cell beyond it along the other axes.
int planeBand(int D, const GridInfo &g, const double *origin, const double *passedNormal, std::vector<int32_t> &idx,
              std::vector<double> &val) {
  double normal[3] = {0, 0, 0}, modulus = 0.;
  for (int i = 0; i < D; ++i) modulus += passedNormal[i] * passedNormal[i];
  modulus = std::sqrt(modulus);
  if (!(modulus > 0)) return -1;
  for (int i = 0; i < D; ++i) normal[i] = passedNormal[i] / modulus;
  int i = 0;
  bool infSet = false;
  for (int n = 0; n < D; ++n)
    if (g.bc[n] == INFINITE) {
      if (infSet) return -2;
      i = n;
      infSet = true;
    }
  if (!infSet) return -4;
  if (passedNormal[i] == 0.) return -3;
  const int j = (i + 1) % D, k = (i + 2) % D;
  const double dx = g.dx;
  double minCoord[2] = {0, 0}, maxCoord[2] = {0, 0};
  for (int n = 0; n < D - 1; ++n) {
    minCoord[n] = dx * (minIndexOf(g, (i + n + 1) % D) - 1);
    maxCoord[n] = dx * (maxIndexOf(g, (i + n + 1) % D) + 1);
  }
  std::vector<Node> corners(2 * (D - 1), Node{0, 0, 0});
  corners[0][j] = minCoord[0];
  corners[1][j] = maxCoord[0];
  if (D == 3) {
    corners[0][k] = minCoord[1];
    corners[1][k] = maxCoord[1];
    corners[2][j] = minCoord[0];
    corners[2][k] = maxCoord[1];
    corners[3][j] = maxCoord[0];
    corners[3][k] = minCoord[1];
  }
  for (auto &c : corners) {
    double numerator = (c[j] - origin[j]) * normal[j];
    if (D == 3) numerator += (c[k] - origin[k]) * normal[k];
    else c[2] = 0.;
    c[i] = origin[i] - numerator / normal[i];
  }
  std::vector<std::array<int, 3>> elements;
  if (D == 2) {
    std::array<int, 3> line = {0, 1, 0};
    if (normal[i] < 0.) std::swap(line[0], line[1]);
    elements.push_back(line);
  } else {
    std::array<int, 3> t = {0, 1, 2};
    if (normal[i] < 0.) std::swap(t[0], t[1]);
    elements.push_back(t);
    t = {0, 3, 1};
    if (normal[i] < 0.) std::swap(t[0], t[1]);
    elements.push_back(t);
  }
  const bool removeAll[3] = {true, true, true};
  fromSurfaceMesh(D, g, corners, elements, removeAll, idx, val);
  return 0;
}

// MakeGeometry::makeBox (lsMakeGeometry.hpp:363-419): the 2^D corners, four
// lines or twelve triangles; boundary triangles are kept along periodic
// axes unless bit d of ignoreBcMask (setIgnoreBoundaryConditions) is set.
void boxBand(int D, const GridInfo &g, const double *mn, const double *mx, std::vector<int32_t> &idx,
             std::vector<double> &val, int ignoreBcMask) {
  std::vector<Node> c(1u << D, Node{0, 0, 0});
  for (int i = 0; i < D; ++i) {
    c[0][i] = mn[i];
    c.back()[i] = mx[i];
  }
  c[1] = c[0];
  c[1][0] = c.back()[0];
  c[2] = c[0];
  c[2][1] = c.back()[1];
  if (D == 3) {
    c[3] = c.back();
    c[3][2] = c[0][2];
    c[4] = c[0];
    c[4][2] = c.back()[2];
    c[5] = c.back();
    c[5][1] = c[0][1];
    c[6] = c.back();
    c[6][0] = c[0][0];
  }
  std::vector<std::array<int, 3>> elements;
  if (D == 2) {
    const std::array<int, 3> lines[4] = {{0, 2, 0}, {2, 3, 0}, {3, 1, 0}, {1, 0, 0}};
    for (const auto &l : lines) elements.push_back(l);
  } else {
    const std::array<int, 3> tri[12] = {{0, 3, 1}, {0, 2, 3}, {0, 1, 5}, {0, 5, 4}, {0, 4, 2}, {4, 6, 2},
                                        {7, 6, 4}, {7, 4, 5}, {7, 2, 6}, {7, 3, 2}, {1, 3, 5}, {3, 7, 5}};
    for (const auto &t : tri) elements.push_back(t);
  }
  const bool removeTri[3] = {(ignoreBcMask & 1) != 0, (ignoreBcMask & 2) != 0, (ignoreBcMask & 4) != 0};
  fromSurfaceMesh(D, g, c, elements, removeTri, idx, val);
}

}  // namespace lsg
