// Per-item logic of the advection kernels: the pointwise rules of the band
// operations, the Hamiltonians, the candidate-bitmap word and the index
// rebuild of one bitmap word.  The kernels in lsg_kernels.hip are grid-stride
// loops around these functions.  They are __host__ __device__ so that the
// same code can be exercised on the host.
//
// Floating-point expressions keep the reference's operation order (files
// cited per function) and the library is built with -ffp-contract=off.
#pragma once

#include "lsg_device.hpp"

namespace lsg {

LSG_HD inline double dmaxs(double a, double b) { return (a < b) ? b : a; }  // std::max
LSG_HD inline double dmins(double a, double b) { return (b < a) ? b : a; }  // std::min
LSG_HD inline double pow2(double v) { return v * v; }

// ------------------------------------------------------------ rules ----

// Expand<T,D>::apply inner rule (lsExpand.hpp:107-157)
template <int D> LSG_HD inline double ruleExpand(double c, const double *nb, double limit, double totalLimit) {
  if (fabs(c) <= totalLimit) return c;
  if (c > -DEPS) {
    double dist = POS;
    LSG_UNROLL
    for (int i = 0; i < 2 * D; i++) {
      double nv = nb[i] + 1.0;
      if (dist > nv) dist = nv;
    }
    return (dist <= limit) ? dist : POS;
  }
  double dist = NEG;
  LSG_UNROLL
  for (int i = 0; i < 2 * D; i++) {
    double nv = nb[i] - 1.0;
    if (dist < nv) dist = nv;
  }
  return (dist >= -limit) ? dist : NEG;
}

// Advect::rebuildLS inner rule (lsAdvect.hpp:304-414)
template <int D> LSG_HD inline double ruleRebuild(double c, const double *nb, double cutoff) {
  if (fabs(c) <= 1.0) {
    int k = 0;
    LSG_UNROLL
    for (; k < 2 * D; k++)
      if (signbit(nb[k] - 1e-7) != signbit(c + 1e-7)) break;
    if (k != 2 * D) {
      if (c > 0.5) {
        int j = 0;
        LSG_UNROLL
        for (; j < 2 * D; j++)
          if (fabs(nb[j]) <= 1.0 && nb[j] < -0.5) break;
        return (j == 2 * D) ? c : 0.5;
      } else if (c < -0.5) {
        int j = 0;
        LSG_UNROLL
        for (; j < 2 * D; j++)
          if (fabs(nb[j]) <= 1.0 && nb[j] > 0.5) break;
        return (j == 2 * D) ? c : -0.5;
      }
      return c;
    }
    return (c < 0) ? NEG : POS;
  }
  if (c >= 0) {
    double dist = POS;
    LSG_UNROLL
    for (int i = 0; i < 2 * D; i++) {
      double v = nb[i];
      if (fabs(v) <= 1.0 && (v < 0.)) {
        if (dist > v + 1.0) dist = v + 1.0;
      }
    }
    return (dist <= cutoff) ? dist : POS;
  }
  double dist = NEG;
  LSG_UNROLL
  for (int i = 0; i < 2 * D; i++) {
    double v = nb[i];
    if (fabs(v) <= 1.0 && (v > 0)) {
      if (dist < v - 1.0) dist = v - 1.0;
    }
  }
  return (dist >= -cutoff) ? dist : NEG;
}

// Prune<T,D>::apply inner rule (lsPrune.hpp:22-52,121-179)
LSG_HD inline bool pIsNeg(double a) { return a <= -DEPS; }
LSG_HD inline bool pSignDiff(double a, double b) { return pIsNeg(a) ^ pIsNeg(b); }
LSG_HD inline bool pSignDiffOrZero(double a, double b) {
  if (a == 0. || b == 0.) return true;
  return pSignDiff(a, b);
}
LSG_HD inline bool pMonotone(double a, double c) { return a == 0. || c == 0. || pSignDiff(a, c); }

template <int D> LSG_HD inline double rulePrune(double c, const double *nb, bool rsz) {
  const bool centerSign = pIsNeg(c);
  if (!isUndef(c)) {
    bool keep = true;
    if (fabs(c) != 0.) {
      keep = false;
      LSG_UNROLL
      for (int i = 0; i < 2 * D; ++i) {
        if (rsz ? pSignDiffOrZero(c, nb[i]) : pSignDiff(c, nb[i])) {
          keep = true;
          break;
        }
      }
    }
    if (rsz && fabs(c) == 0.) {
      LSG_UNROLL
      for (int i = 0; i < D; ++i) {
        double negVal = nb[i];
        double posVal = nb[D + i];
        if (!pMonotone(negVal, posVal)) return fabs(negVal) < fabs(posVal) ? negVal : posVal;
      }
    }
    if (keep) return c;
  }
  return centerSign ? NEG : POS;
}

// BooleanOperation comparators (lsBooleanOperation.hpp:250-269) and the
// CUSTOM comparator of Advect::combineLevelSets (lsAdvect.hpp:218-230);
// undefined results keep their sign (lsBooleanOperation.hpp:115-118).
LSG_HD inline double ruleBoolean(int kind, double a, double b, double wT, double wS) {
  double r;
  if (kind == B_INTERSECT) r = (a > b) ? a : b;
  else if (kind == B_UNION) r = (a < b) ? a : b;
  else if (kind == B_RELCOMP) r = (a > -b) ? a : -b;
  else {
    if (!isUndef(a) && !isUndef(b)) r = wS * a + wT * b;
    else if (isUndef(a)) r = a;
    else r = b;
  }
  if (isUndef(r)) r = (r < 0) ? NEG : POS;
  return r;
}

// Reduce rule (lsReduce.hpp:106-118) and updateLevelSet for one level set
// (lsAdvect.hpp:700,725-762).
LSG_HD inline double ruleReduce(double c, double limit) {
  return (!isUndef(c) && fabs(c) <= limit) ? c : ((c < 0) ? NEG : POS);
}
LSG_HD inline double ruleUpdate(double c, double dt, double velocity) {
  if (!isUndef(c) && fabs(c) <= 0.5) {
    double rate = dt * velocity;
    return c - rate;
  }
  return (c < 0) ? NEG : POS;
}
// updateLevelSet with several level sets (lsAdvect.hpp:725-771): the point's
// rate list holds (velocity after the dissipation rule, stop value) pairs;
// while the step would carry the value past a stop value, the value moves
// to it and the remaining time continues with the next entry.
LSG_HD inline double ruleUpdateMM(double c, double dt, const double *list) {
  if (!isUndef(c) && fabs(c) <= 0.5) {
    double value = c, time = dt;
    int k = 0;
    double velocity = list[0];
    double rate = time * velocity;
    while (fabs(list[2 * k + 1] - value) < fabs(rate)) {
      time -= fabs((list[2 * k + 1] - value) / velocity);
      value = list[2 * k + 1];
      ++k;
      velocity = list[2 * k];
      rate = time * velocity;
    }
    return value - rate;
  }
  return (c < 0) ? NEG : POS;
}

// ------------------------------------------------- point-data sources ----
// newDataSourceIds of the band operations: the point whose data a defined
// result takes -- 2D for the centre, k for star neighbour k (nb order), -1
// for none (no id is recorded).

enum SrcKind : int { SRC_CENTRE = 0, SRC_EXPAND = 1, SRC_REBUILD = 2, SRC_PRUNE = 3, SRC_PRUNE_RSZ = 4, SRC_BOOLEAN = 5 };

// Expand (lsExpand.hpp:107-157): the centre, else the neighbour that set the
// distance (first of equal ones).
template <int D> LSG_HD inline int srcExpand(double c, const double *nb, double totalLimit) {
  if (fabs(c) <= totalLimit) return 2 * D;
  int k = -1;
  if (c > -DEPS) {
    double dist = POS;
    LSG_UNROLL
    for (int i = 0; i < 2 * D; ++i)
      if (dist > nb[i] + 1.0) {
        dist = nb[i] + 1.0;
        k = i;
      }
  } else {
    double dist = NEG;
    LSG_UNROLL
    for (int i = 0; i < 2 * D; ++i)
      if (dist < nb[i] - 1.0) {
        dist = nb[i] - 1.0;
        k = i;
      }
  }
  return k;
}
// Advect::rebuildLS (lsAdvect.hpp:304-413)
template <int D> LSG_HD inline int srcRebuild(double c, const double *nb) {
  if (fabs(c) <= 1.0) {
    if (c > 0.5) {
      for (int j = 0; j < 2 * D; j++)
        if (fabs(nb[j]) <= 1.0 && nb[j] < -0.5) return j;
    } else if (c < -0.5) {
      for (int j = 0; j < 2 * D; j++)
        if (fabs(nb[j]) <= 1.0 && nb[j] > 0.5) return j;
    }
    return 2 * D;
  }
  int used = -1;
  if (c >= 0) {
    double dist = POS;
    for (int i = 0; i < 2 * D; i++)
      if (fabs(nb[i]) <= 1.0 && (nb[i] < 0.) && dist > nb[i] + 1.0) {
        dist = nb[i] + 1.0;
        used = i;
      }
  } else {
    double dist = NEG;
    for (int i = 0; i < 2 * D; i++)
      if (fabs(nb[i]) <= 1.0 && (nb[i] > 0) && dist < nb[i] - 1.0) {
        dist = nb[i] - 1.0;
        used = i;
      }
  }
  return used;
}
// Prune (lsPrune.hpp:135-167): a kept point records itself; the stray-zero
// rewrite records nothing.
template <int D> LSG_HD inline int srcPrune(double c, const double *nb, bool rsz) {
  if (rsz && fabs(c) == 0.) {
    LSG_UNROLL
    for (int i = 0; i < D; ++i)
      if (!pMonotone(nb[i], nb[D + i])) return -1;
  }
  return 2 * D;
}
// The comparator's second member: the result is taken from A
// (lsBooleanOperation.hpp:250-280, lsAdvect.hpp:218-230).
LSG_HD inline bool boolFromA(int kind, double a, double b) {
  if (kind == B_INTERSECT) return a > b;
  if (kind == B_UNION) return a < b;
  if (kind == B_RELCOMP) return a > -b;
  return !isUndef(a) && !isUndef(b);
}

// The rate (time * velocity) of the last list entry a point's update walks
// to, and that entry's index (saveAdvectionVelocities, lsAdvect.hpp:733-765).
LSG_HD inline double updateRateMM(double c, double dt, const double *list, int &last) {
  double value = c, time = dt;
  int k = 0;
  double velocity = list[0];
  double rate = time * velocity;
  while (fabs(list[2 * k + 1] - value) < fabs(rate)) {
    time -= fabs((list[2 * k + 1] - value) / velocity);
    value = list[2 * k + 1];
    ++k;
    velocity = list[2 * k];
    rate = time * velocity;
  }
  last = k;
  return rate;
}

// ------------------------------------------------------- neighbourhood ----

// Frame bit index of grid point c and, per axis d and step k (|k| <= R),
// the frame-bit offset off[d][R + k] of map(d, c_d + k) from c, when every
// such point is a frame bit: c inside the grid, each bounded axis (which
// spans the frame) at least R + 1 points long, so that one wrap or mirror
// is the whole mapping of a step, and c at least R cells inside the frame
// along each unbounded axis, where the mapping is the identity.  -1
// otherwise.  A located point is then exactly that index with inside =
// true, and a boundary costs a few selects instead of locate() and map().
template <int D> LSG_HD inline int64_t axisStride(const BandView &b, int d);
template <int D, int R> LSG_HD inline int64_t mappedLin(const BandView &b, const int32_t c[3], int64_t off[][2 * R + 1]) {
  bool ok = true;
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {
    const int64_t lo = b.g.lo[d], hi = b.g.hi[d];
    const int64_t st = axisStride<D>(b, d);
    if (b.g.bounded(d)) {
      ok = ok && c[d] >= lo && c[d] <= hi && hi - lo >= R;
      const bool per = b.g.bc[d] == PERIODIC;
      const int64_t L = hi - lo + 1;
      LSG_UNROLL
      for (int k = -R; k <= R; ++k) {
        const int64_t y = (int64_t)c[d] + k;
        int64_t m = y;
        if (k < 0 && y < lo) m = per ? y + L : 2 * lo - y;
        if (k > 0 && y > hi) m = per ? y - L : 2 * hi - y;
        off[d][R + k] = (m - c[d]) * st;
      }
    } else {
      const int64_t r = (int64_t)c[d] - b.f.o[d];
      ok = ok && r >= R && r + R < b.f.ext[d];
      LSG_UNROLL
      for (int k = -R; k <= R; ++k) off[d][R + k] = k * st;
    }
  }
  const int64_t xr = (int64_t)c[0] - b.f.o[0];
  const int64_t yr = (int64_t)c[1] - b.f.o[1];
  const int64_t zr = (int64_t)c[2] - b.f.o[2];
  if (D == 2) ok = ok && zr >= 0 && zr < b.f.ext[2];
  return ok ? ((zr * b.f.ext[1] + yr) * (int64_t)b.f.rowBits() + xr) : -1;
}
LSG_HD inline BandView::Loc locAt(int64_t lin) {
  BandView::Loc l;
  l.w = (uint32_t)(lin >> 6);
  l.bt = (uint32_t)(lin & 63);
  l.inside = true;
  return l;
}
template <int D> LSG_HD inline int64_t axisStride(const BandView &b, int d) {
  return d == 0 ? 1 : (d == 1 ? (int64_t)b.f.rowBits() : (int64_t)b.f.ext[1] * b.f.rowBits());
}

// Centre value and the 2D star neighbours (boundary mapped) of point c:
// nb[d] = +e_d, nb[D + d] = -e_d.  All 2D + 1 lookups are issued together.
template <int D> LSG_HD inline void starLocs(const BandView &b, const int32_t c[3], BandView::Loc *l) {
  int64_t off[3][3];
  const int64_t lin = mappedLin<D, 1>(b, c, off);
  if (lin >= 0) {
    l[2 * D] = locAt(lin);
    LSG_UNROLL
    for (int d = 0; d < D; ++d) {
      l[d] = locAt(lin + off[d][2]);
      l[D + d] = locAt(lin + off[d][0]);
    }
    return;
  }
  l[2 * D] = b.locate(c[0], c[1], c[2]);
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {
    int32_t p[3] = {c[0], c[1], c[2]};
    p[d] = b.g.map(d, (int64_t)c[d] + 1);
    l[d] = b.locate(p[0], p[1], p[2]);
    int32_t q[3] = {c[0], c[1], c[2]};
    q[d] = b.g.map(d, (int64_t)c[d] - 1);
    l[D + d] = b.locate(q[0], q[1], q[2]);
  }
}
template <int D> LSG_HD inline void gatherStar(const BandView &b, const int32_t c[3], double &center, double *nb) {
  BandView::Loc l[2 * D + 1];
  starLocs<D>(b, c, l);
  double v[2 * D + 1];
  b.values<2 * D + 1>(l, v);
  LSG_UNROLL
  for (int k = 0; k < 2 * D; ++k) nb[k] = v[k];
  center = v[2 * D];
}
// Values along every axis of point c up to distance R (boundary mapped):
// st[d][R + k] = phi(c + k e_d) for 0 < |k| <= R; st[d][R] is left to the
// caller (the centre value).  All 2 D R lookups are issued together.
template <int D, int R> LSG_HD inline void gatherAxes(const BandView &b, const int32_t c[3], double st[][2 * R + 1]) {
  BandView::Loc l[2 * D * R];
  int64_t off[3][2 * R + 1];
  const int64_t lin = mappedLin<D, R>(b, c, off);
  if (lin >= 0) {
    LSG_UNROLL
    for (int d = 0; d < D; ++d) {
      LSG_UNROLL
      for (int k = 1; k <= R; ++k) {
        l[(d * R + k - 1) * 2] = locAt(lin + off[d][R + k]);
        l[(d * R + k - 1) * 2 + 1] = locAt(lin + off[d][R - k]);
      }
    }
  } else {
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {
    LSG_UNROLL
    for (int k = 1; k <= R; ++k) {
      int32_t p[3] = {c[0], c[1], c[2]};
      p[d] = b.g.map(d, (int64_t)c[d] + k);
      l[(d * R + k - 1) * 2] = b.locate(p[0], p[1], p[2]);
      int32_t q[3] = {c[0], c[1], c[2]};
      q[d] = b.g.map(d, (int64_t)c[d] - k);
      l[(d * R + k - 1) * 2 + 1] = b.locate(q[0], q[1], q[2]);
    }
  }
  }
  double v[2 * D * R];
  b.values<2 * D * R>(l, v);
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {
    LSG_UNROLL
    for (int k = 1; k <= R; ++k) {
      st[d][R + k] = v[(d * R + k - 1) * 2];
      st[d][R - k] = v[(d * R + k - 1) * 2 + 1];
    }
  }
}

// updateLevelSet followed by rebuildLS (ruleUpdate, then ruleRebuild with
// the given cutoff) at point c: the rebuild rule on the updated values of
// c's star, each the update rule of the input entry there and its rate
// (rates are indexed by input entry, like the update operation's).  A
// point that is not a candidate of the rebuild sees undefined values only
// (the update keeps an undefined entry's sign), so it gets the value its
// run holds in the rebuilt band.
template <int D>
LSG_HD inline double updateRebuild(const BandView &b, const int32_t c[3], const double *rates, double dt,
                                   double cutoff) {
  constexpr int N = 2 * D + 1;
  BandView::Loc l[N];
  starLocs<D>(b, c, l);
  uint64_t wd[N];
  uint32_t pr[N];
  LSG_UNROLL
  for (int k = 0; k < N; ++k) {
    wd[k] = b.bits[l[k].w];
    pr[k] = b.pre[l[k].w];
  }
  double u[N];
  LSG_UNROLL
  for (int k = 0; k < N; ++k) {
    const uint32_t e = b.entryOf(l[k], wd[k], pr[k]);
    u[k] = ruleUpdate(b.val[e], dt, rates[e]);
  }
  return ruleRebuild<D>(u[2 * D], u, cutoff);
}

// ruleExpand from the minimum and the maximum of the star's values: the
// rule's minimum of (neighbour + 1) started at POS is (minimum + 1) and its
// maximum of (neighbour - 1) started at NEG is (maximum - 1), exactly, since
// rounding is monotone (POS + 1 rounds to POS).
LSG_HD inline double ruleExpandMinMax(double c, double vmin, double vmax, double limit, double totalLimit) {
  if (fabs(c) <= totalLimit) return c;
  if (c > -DEPS) {
    const double dist = vmin + 1.0;
    return (dist <= limit) ? dist : POS;
  }
  const double dist = vmax - 1.0;
  return (dist >= -limit) ? dist : NEG;
}
LSG_HD inline void minMaxAcc(double v, double &vmin, double &vmax) {
  vmin = (v < vmin) ? v : vmin;
  vmax = (v > vmax) ? v : vmax;
}

// Two Expand cycles at once (ruleExpand with limit lim1, then lim2; final
// limit tot) at point c of the grid, from the input band within two star
// steps of c: the first cycle's rule at c and at each of its boundary-mapped
// star neighbours, then the second cycle's rule at c on those values -- the
// composition of two cycles.  A point that is not a candidate of the first
// cycle sees undefined values only, so the rule gives it the value its run
// holds in the first cycle's result.  The steps are boundary mapped one at
// a time (a mirror boundary makes two steps along one axis depend on the
// path): m1[d][s] = map(c_d +- 1) (s = 0: +, 1: -) is a star point's
// coordinate along d, m2[d][s][t] = map(m1[d][s] +- 1) the steps along d
// from it (the centre and c +- 2 e_d away from boundaries), and the
// diagonals have both axes at their m1 (the same from either star point:
// the axes map independently).  Bounded axes span the frame, so wherever c
// lies two cells inside the frame along the unbounded axes every point is
// a frame bit at a fixed offset from c; elsewhere each point is located.
// The points are read in two batches to bound the loads in flight.
template <int D>
LSG_HD inline double expandTwice(const BandView &b, const int32_t c[3], double lim1, double lim2, double tot) {
  int32_t m1[3][2], m2[3][2][2];
  int64_t o1[3][2], o2[3][2][2];  // frame-bit offsets of the mapped coordinates from c
  bool fast = true;
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {
#ifdef LSG_HOST_ONLY
    if (c[d] < b.g.lo[d] || c[d] > b.g.hi[d]) std::abort();  // the host build checks that c is a grid point
#endif
    int32_t below, above;
    b.g.stepTargets(d, below, above);
    const int64_t st = axisStride<D>(b, d);
    LSG_UNROLL
    for (int s = 0; s < 2; ++s) {
      m1[d][s] = GridInfo::stepTo((int64_t)c[d] + (s ? -1 : 1), b.g.lo[d], b.g.hi[d], below, above);
      o1[d][s] = (int64_t)(m1[d][s] - c[d]) * st;
      LSG_UNROLL
      for (int t = 0; t < 2; ++t) {
        m2[d][s][t] = GridInfo::stepTo((int64_t)m1[d][s] + (t ? -1 : 1), b.g.lo[d], b.g.hi[d], below, above);
        o2[d][s][t] = (int64_t)(m2[d][s][t] - c[d]) * st;
      }
    }
    if (!b.g.bounded(d)) {
      const int64_t r = (int64_t)c[d] - b.f.o[d];
      fast = fast && r >= 2 && r + 2 < b.f.ext[d];
    }
  }
  const int64_t lin0 = ((int64_t)(c[2] - b.f.o[2]) * b.f.ext[1] + (c[1] - b.f.o[1])) * (int64_t)b.f.rowBits() +
                       (c[0] - b.f.o[0]);
  // batch 1: slot 0 the centre, 1 + 2d + s the star points, then per axis
  // pair lo < hi the four diagonals (lo at m1[lo][k >> 1], hi at
  // m1[hi][k & 1])
  constexpr int N1 = 1 + 2 * D + 2 * D * (D - 1);
  double v1c, qm[3][2], qx[3][2], sv[3][2];
  {
    BandView::Loc l[N1];
    if (fast) {
      l[0] = locAt(lin0);
      LSG_UNROLL
      for (int d = 0; d < D; ++d)
        LSG_UNROLL
        for (int s = 0; s < 2; ++s) l[1 + 2 * d + s] = locAt(lin0 + o1[d][s]);
      LSG_UNROLL
      for (int lo = 0; lo < D; ++lo)
        LSG_UNROLL
        for (int hi = lo + 1; hi < D; ++hi) {
          const int q = 1 + 2 * D + 4 * (lo * (2 * D - lo - 1) / 2 + (hi - lo - 1));
          LSG_UNROLL
          for (int k = 0; k < 4; ++k) l[q + k] = locAt(lin0 + o1[lo][k >> 1] + o1[hi][k & 1]);
        }
    } else {
      l[0] = b.locate(c[0], c[1], c[2]);
      LSG_UNROLL
      for (int d = 0; d < D; ++d)
        LSG_UNROLL
        for (int s = 0; s < 2; ++s) {
          int32_t p[3] = {c[0], c[1], c[2]};
          p[d] = m1[d][s];
          l[1 + 2 * d + s] = b.locate(p[0], p[1], p[2]);
        }
      LSG_UNROLL
      for (int lo = 0; lo < D; ++lo)
        LSG_UNROLL
        for (int hi = lo + 1; hi < D; ++hi) {
          const int q = 1 + 2 * D + 4 * (lo * (2 * D - lo - 1) / 2 + (hi - lo - 1));
          LSG_UNROLL
          for (int k = 0; k < 4; ++k) {
            int32_t p[3] = {c[0], c[1], c[2]};
            p[lo] = m1[lo][k >> 1];
            p[hi] = m1[hi][k & 1];
            l[q + k] = b.locate(p[0], p[1], p[2]);
          }
        }
    }
    double v[N1];
    b.values<N1>(l, v);
    double mn = POS, mx = NEG;
    LSG_UNROLL
    for (int k = 0; k < 2 * D; ++k) minMaxAcc(v[1 + k], mn, mx);
    v1c = ruleExpandMinMax(v[0], mn, mx, lim1, tot);
    LSG_UNROLL
    for (int a = 0; a < D; ++a)
      LSG_UNROLL
      for (int s = 0; s < 2; ++s) {
        sv[a][s] = v[1 + 2 * a + s];
        qm[a][s] = POS;
        qx[a][s] = NEG;
        LSG_UNROLL
        for (int d = 0; d < D; ++d) {
          if (d == a) continue;
          const int lo = a < d ? a : d, hi = a < d ? d : a;
          const int q = 1 + 2 * D + 4 * (lo * (2 * D - lo - 1) / 2 + (hi - lo - 1));
          // axis a fixed at m1[a][s], axis d at both of its m1
          if (a < d) {
            minMaxAcc(v[q + 2 * s], qm[a][s], qx[a][s]);
            minMaxAcc(v[q + 2 * s + 1], qm[a][s], qx[a][s]);
          } else {
            minMaxAcc(v[q + s], qm[a][s], qx[a][s]);
            minMaxAcc(v[q + 2 + s], qm[a][s], qx[a][s]);
          }
        }
      }
  }
  // batch 2: slot 4d + 2s + t the point with axis d at m2[d][s][t]
  double mn2 = POS, mx2 = NEG;
  {
    BandView::Loc l[4 * D];
    LSG_UNROLL
    for (int d = 0; d < D; ++d)
      LSG_UNROLL
      for (int s = 0; s < 2; ++s)
        LSG_UNROLL
        for (int t = 0; t < 2; ++t) {
          if (fast) {
            l[4 * d + 2 * s + t] = locAt(lin0 + o2[d][s][t]);
          } else {
            int32_t p[3] = {c[0], c[1], c[2]};
            p[d] = m2[d][s][t];
            l[4 * d + 2 * s + t] = b.locate(p[0], p[1], p[2]);
          }
        }
    double w[4 * D];
    b.values<4 * D>(l, w);
    LSG_UNROLL
    for (int a = 0; a < D; ++a)
      LSG_UNROLL
      for (int s = 0; s < 2; ++s) {
        minMaxAcc(w[4 * a + 2 * s], qm[a][s], qx[a][s]);
        minMaxAcc(w[4 * a + 2 * s + 1], qm[a][s], qx[a][s]);
        minMaxAcc(ruleExpandMinMax(sv[a][s], qm[a][s], qx[a][s], lim1, tot), mn2, mx2);
      }
  }
  return ruleExpandMinMax(v1c, mn2, mx2, lim2, tot);
}

// Coordinates of candidate / entry i at frame position p (entry 0 is the
// grid minimum point when it lies outside the frame).
LSG_HD inline void candCoords(const BandView &b, uint32_t i, uint32_t p, int32_t c[3]) {
  (void)i;
  if (p == POS_NONE) {
    c[0] = b.g.lo[0];
    c[1] = b.g.lo[1];
    c[2] = b.g.lo[2];
  } else {
    b.coords(p, c);
  }
}

// Star-dilation candidate word (visit set of HRLE's
// ConstSparseStarIterator<.,1>): grid points whose boundary-mapped star
// neighbourhood holds a run start, their lexicographic successors, and the
// first point of every row when x is bounded (successor of the previous
// row's last point).
template <int D> LSG_HD inline uint64_t candStarWord(const uint64_t *bits, const Frame &f, const GridInfo &g, int64_t w) {
  const int wpr = f.wpr;
  const int lastBit = f.ext[0] - 1;
  const int lastWord = lastBit >> 6;
  int32_t kw, yr, zr;
  f.wordRow((uint32_t)w, kw, yr, zr);
  const uint64_t B = bits[w];
  const uint64_t BL = kw > 0 ? bits[w - 1] : 0;
  const uint64_t BH = kw + 1 < wpr ? bits[w + 1] : 0;
  uint64_t dil = B | (B << 1) | (BL >> 63) | (B >> 1) | (BH << 63);
  uint64_t prevTop = 0;
  if (kw > 0) {
    const uint64_t BLL = kw > 1 ? bits[w - 2] : 0;
    const uint64_t dl = BL | (BL << 1) | (BLL >> 63) | (BL >> 1) | (B << 63);
    prevTop = dl >> 63;
  }
  if (g.bc[0] == PERIODIC && kw == lastWord) {
    const uint64_t first = bits[w - kw];  // x = xmin - 1 wraps to xmax
    if (first & 1ull) dil |= 1ull << (lastBit & 63);
  }
  const int32_t ca[3] = {0, yr + f.o[1], zr + f.o[2]};
  LSG_UNROLL
  for (int d = 1; d < D; ++d) {
    LSG_UNROLL
    for (int s = -1; s <= 1; s += 2) {
      int32_t m = g.map(d, (int64_t)ca[d] + s);
      int32_t rr = m - f.o[d];
      if (rr < 0 || rr >= f.ext[d]) continue;
      int32_t ny = (d == 1) ? rr : yr;
      int32_t nz = (d == 2) ? rr : zr;
      int64_t nrow = (int64_t)nz * f.ext[1] + ny;
      int64_t nwi = nrow * wpr + kw;
      dil |= bits[nwi];
      if (kw > 0) prevTop |= bits[nwi - 1] >> 63;
    }
  }
  uint64_t C = dil | (dil << 1) | prevTop;
  if (g.bounded(0) && kw == 0) C |= 1ull;
  if (kw == lastWord) {
    const int nb = (lastBit & 63) + 1;
    if (nb < 64) C &= (~0ull >> (64 - nb));
  }
  return C;
}

// Parallel bit deposit: the low popcount(m) bits of x placed, in order, at
// the set bits of m (Hacker's Delight, 2nd ed., 7-5 "expand", 64-bit).
LSG_HD inline uint64_t depositBits(uint64_t x, uint64_t m) {
  const uint64_t m0 = m;
  uint64_t mk = ~m << 1;
  uint64_t mvs[6];
  LSG_UNROLL
  for (int i = 0; i < 6; ++i) {
    uint64_t mp = mk ^ (mk << 1);
    mp ^= mp << 2;
    mp ^= mp << 4;
    mp ^= mp << 8;
    mp ^= mp << 16;
    mp ^= mp << 32;
    const uint64_t mv = mp & m;
    mvs[i] = mv;
    m = (m ^ mv) | (mv >> (1 << i));
    mk &= ~mp;
  }
  LSG_UNROLL
  for (int i = 5; i >= 0; --i) {
    const uint64_t mv = mvs[i];
    const uint64_t t = x << (1 << i);
    x = (x & ~mv) | (t & mv);
  }
  return x & m0;
}

// Index of the k-th (0-based) set bit of w (k < popcount(w)).
LSG_HD inline int selectBit(uint64_t w, int k) {
  int pos = 0;
  LSG_UNROLL
  for (int sh = 32; sh >= 1; sh >>= 1) {
    const uint64_t low = w & ((sh == 64) ? ~0ull : ((1ull << sh) - 1ull));
    const int c = __builtin_popcountll(low);
    if (k >= c) {
      k -= c;
      w >>= sh;
      pos += sh;
    } else {
      w = low;
    }
  }
  return pos;
}

// Frame position of the grid point before in-frame position p (p > 0) in
// HRLE order: the previous bit, or the last x of the previous row when p
// starts a row (rows are padded to whole words, the padding never holds a
// point).
LSG_HD inline uint32_t realPred(const Frame &f, uint32_t p) {
  const uint32_t w = p >> 6;
  if ((p & 63u) != 0u) return p - 1u;
  const uint32_t row = f.divW.div(w);
  if (w != row * (uint32_t)f.wpr) return p - 1u;  // bit 63 of the previous word of the row
  return ((w - 1u) << 6) | (uint32_t)((f.ext[0] - 1) & 63);
}

// Largest j < N with pre[j] <= c for a nondecreasing pre with pre[0] <= c:
// the word of a tile that holds item c of the tile (N a power of two).
template <int N> LSG_HD inline int wordOfItem(const uint32_t *pre, uint32_t c) {
  int lo = 0;
  LSG_UNROLL
  for (int step = N / 2; step >= 1; step >>= 1)
    if (pre[lo + step] <= c) lo += step;
  return lo;
}

// Slot s of tile t (words w0.. of the frame, candidate words cw with
// exclusive prefix counts cpre) of a single-pass band operation: slot 0 is
// the grid point before the tile's first candidate (for tile 0 the
// grid-minimum entry outside the frame, when there is one), slot s >= 1 is
// candidate s - 1.  Sets the frame position p (POS_NONE for the grid-minimum
// entry), whether the rule is evaluated there and whether it is a candidate
// (word j, bit b of the tile).
template <int N>
LSG_HD inline void tileSlot(const Frame &f, uint32_t t, uint32_t w0, const uint64_t *cw, const uint32_t *cpre,
                            uint32_t nc, uint32_t s, uint32_t &p, bool &eval, bool &cand, int &j, int &b) {
  p = POS_NONE;
  eval = cand = false;
  j = b = 0;
  if (s == 0) {
    if (t > 0 || f.base) {
      eval = true;
      if (t > 0) {
        const int j0 = wordOfItem<N>(cpre, 0);
        p = realPred(f, ((w0 + (uint32_t)j0) << 6) | (uint32_t)__builtin_ctzll(cw[j0]));
      }
    }
  } else if (s <= nc) {
    const uint32_t c = s - 1;
    j = wordOfItem<N>(cpre, c);
    b = selectBit(cw[j], (int)(c - cpre[j]));
    p = ((w0 + (uint32_t)j) << 6) | (uint32_t)b;
    eval = cand = true;
  }
}

// Rules of the band operations evaluated at a frame position p (POS_NONE:
// the grid-minimum entry outside the frame).  Evaluated at a point that is
// not a candidate of the operation they return the value of the result's
// run holding that point, which is what the keep flag of the next candidate
// compares against: such a point's boundary-mapped star holds no run start,
// so the rule sees one undefined value everywhere and returns it (star
// rules), and a self rule evaluated "at or before" p reads the entry whose
// run holds p.
template <int D> struct AtStar {
  BandView in;
  OpParams op;
  LSG_HD double operator()(uint32_t p) const {
    int32_t c[3];
    candCoords(in, 0, p, c);
    double center, nb[2 * D];
    gatherStar<D>(in, c, center, nb);
    return (op.kind == OP_EXPAND) ? ruleExpand<D>(center, nb, op.a, op.b) : ruleRebuild<D>(center, nb, op.a);
  }
};

template <int D> struct AtSelf {
  BandView in;
  OpParams op;
  LSG_HD double operator()(uint32_t p) const {
    // the grid-minimum entry precedes the band's first in-frame entry (in a
    // batch view: the level set's own, after the entries of earlier sets)
    uint32_t i = in.rb + in.pre[0] - 1;
    if (p != POS_NONE) {
      const BandView::Loc l = locAt((int64_t)p);
      i = in.entryOf(l, in.bits[l.w], in.pre[l.w]);
    }
    const double c = in.val[i];
    if (op.kind == OP_REDUCE) return ruleReduce(c, op.a);
    if (op.kind == OP_UPDATE) return ruleUpdate(c, *op.dt, op.rates[i]);
    if (op.kind == OP_UPDATE_MM) return ruleUpdateMM(c, *op.dt, op.rates + (size_t)i * 2 * op.nmat);
    int32_t cc[3];
    candCoords(in, i, in.pos[i], cc);
    double center, nb[2 * D];
    gatherStar<D>(in, cc, center, nb);
    return rulePrune<D>(c, nb, op.kind == OP_PRUNE_RSZ);
  }
};

struct AtBool {
  BandView A, B;
  int kind;
  double wT, wS;
  uint32_t *movedDown;  // combine only: set where the result exceeds b (both defined)
  LSG_HD double operator()(uint32_t p) const {
    int32_t c[3];
    candCoords(A, 0, p, c);
    BandView::Loc l[2] = {A.locate(c[0], c[1], c[2]), B.locate(c[0], c[1], c[2])};
    const uint64_t wa = A.bits[l[0].w], wb = B.bits[l[1].w];
    const uint32_t pa = A.pre[l[0].w], pb = B.pre[l[1].w];
    const double a = A.val[A.entryOf(l[0], wa, pa)];
    const double b = B.val[B.entryOf(l[1], wb, pb)];
    const double r = ruleBoolean(kind, a, b, wT, wS);
    if (movedDown && kind == B_COMBINE && !isUndef(a) && !isUndef(b) && r > b) *movedDown = 1u;
    return r;
  }
};

// Keep flag of candidate i: first entry, defined, or start of a run whose
// value differs from the previous candidate's.
LSG_HD inline uint32_t keepFlag(const double *out, uint32_t i) {
  double v = out[i];
  if (i == 0) return 1;
  if (!isUndef(v)) return 1;
  return v != out[i - 1] ? 1u : 0u;
}

// Bitmap word and prefix count of the compacted output for word w.
LSG_HD inline void buildIndexWord(const uint32_t *cpre, const uint32_t *ks, const uint32_t *cpos, int64_t w, int64_t nw,
                                  int base, uint32_t ncand, uint64_t &word, uint32_t &pre) {
  uint32_t c0 = (uint32_t)base + cpre[w];
  if (c0 > ncand) c0 = ncand;
  uint32_t k0 = ks[c0];
  pre = k0 - (uint32_t)base;
  word = 0;
  if (w == nw) return;
  uint32_t c1 = (uint32_t)base + cpre[w + 1];
  if (c1 > ncand) c1 = ncand;
  uint32_t kprev = k0;
  for (uint32_t c = c0; c < c1; ++c) {
    uint32_t kn = ks[c + 1];
    if (kn != kprev) word |= 1ull << (cpos[c] & 63);
    kprev = kn;
  }
}

// ------------------------------------------------------------- schemes ----

struct Vel {
  int kind, dir, k;
  double b, aPos, aNeg;
  double vec[3];
  double alpha[3];  // getDissipationAlpha (LLFA): constant per direction
  // host-evaluated velocities at the 3^D points of the local Lax-Friedrichs
  // alpha box: box point q has (scalar, vector[3]) at box[q * boxStride];
  // for the analytical scheme the host's getDissipationAlpha values per
  // direction at box[q * boxStride + 1 + direction]
  const double *box = nullptr;
  int boxStride = 4;
  // stencil local Lax-Friedrichs: the host's getScalarVelocity at the
  // perturbed normals of box point q, axis k: sllf[(q * D + k) * 2] at
  // normal - DN e_k, sllf[(q * D + k) * 2 + 1] at normal + DN e_k
  const double *sllf = nullptr;
  // built-in VelocityField::getScalarVelocity (lsVelocityField.hpp:17-21)
  LSG_HD __attribute__((always_inline)) double scalar(const double n[3]) const {
    if (kind == 0) return b;
    if (kind == 2) {  // b + aPos * sum |n_i| (examples/AirGapDeposition/AirGapDeposition.cpp:20-27)
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += fabs(n[i]);
      return b + aPos * s;
    }
    double c = n[dir];
    double a = (c > 0) ? aPos : aNeg;
    double ac = fabs(c);
    double m = 1.0;
    if (k >= 1) {
      m = ac;
      LSG_UNROLL
      for (int j = 1; j < k; ++j) m = m * ac;
    }
    return b + a * m;
  }
};

// Functor of material m (several level sets): its own if set, else `base`.
LSG_HD inline Vel velOfMaterial(const RateParams &rp, const Vel &base, int m) {
  if (m < 0 || m >= 8 || !rp.matSet[m]) return base;
  Vel v;
  v.kind = rp.matKind[m];
  v.dir = rp.matDir[m];
  v.k = rp.matK[m];
  v.b = rp.matB[m];
  v.aPos = rp.matAPos[m];
  v.aNeg = rp.matANeg[m];
  v.vec[0] = rp.matVec[m][0];
  v.vec[1] = rp.matVec[m][1];
  v.vec[2] = rp.matVec[m][2];
  v.alpha[0] = rp.matAlpha[m][0];
  v.alpha[1] = rp.matAlpha[m][1];
  v.alpha[2] = rp.matAlpha[m][2];
  return v;
}

LSG_HD inline Vel velOf(const RateParams &rp) {
  Vel v;
  v.kind = rp.velKind;
  v.dir = rp.velDir;
  v.k = rp.velK;
  v.b = rp.velB;
  v.aPos = rp.velAPos;
  v.aNeg = rp.velANeg;
  v.vec[0] = rp.vec[0];
  v.vec[1] = rp.vec[1];
  v.vec[2] = rp.vec[2];
  v.alpha[0] = rp.velAlpha[0];
  v.alpha[1] = rp.velAlpha[1];
  v.alpha[2] = rp.velAlpha[2];
  return v;
}

// One level set of a batch: its frame and grid, where its bitmap words
// start in the batch bitmap, the grid-minimum entries of the level sets up
// to and including it (they precede each level set's in-frame entries), and
// its velocity.
struct SegInfo {
  Frame f;
  GridInfo g;
  uint32_t wordOff;
  uint32_t baseIncl;
  Vel vel;
};

// Batch context of a kernel (unused when the kernel runs on one level set).
struct SegArgs {
  const SegInfo *segs;
  uint32_t nseg;
  uint32_t totalBase;     // grid-minimum entries of all level sets
  const uint16_t *wseg;   // level set of each bitmap word
  const uint16_t *iseg;   // level set of each input item (candidate or entry)
  uint16_t *oseg;         // level set of each output entry
};

// View of level set k of a batch band.
LSG_HD inline BandView segView(const BandView &all, const SegInfo &S) {
  BandView v = all;
  v.bits = all.bits + S.wordOff;
  v.pre = all.pre + S.wordOff;
  v.f = S.f;
  v.g = S.g;
  v.rb = S.baseIncl;
  return v;
}

// One-sided differences with the second-order ENO correction
// (lsEngquistOsher.hpp:59-111; identical in lsLaxFriedrichs.hpp:66-119,
// lsLocalLaxFriedrichs.hpp:85-143, lsLocalLocalLaxFriedrichs.hpp:62-115).
LSG_HD inline void upwindDiffs(const double *st, double gridDelta, int order, double phi0, double &diffPos,
                               double &diffNeg) {
  // st: values at offsets -2..2 along the axis (st[2] unused)
  const double deltaPos = gridDelta;
  const double deltaNeg = -gridDelta;
  const double phiPos = st[3];
  const double phiNeg = st[1];
  diffPos = (phiPos - phi0) / deltaPos;
  diffNeg = (phiNeg - phi0) / deltaNeg;
  if (order == 2) {
    const double deltaPosPos = 2 * gridDelta;
    const double deltaNegNeg = -2 * gridDelta;
    const double phiPosPos = st[4];
    const double phiNegNeg = st[0];
    const double diff00 = (((deltaNeg * phiPos - deltaPos * phiNeg) / (deltaPos - deltaNeg) + phi0)) / (deltaPos * deltaNeg);
    const double diffNegNeg =
        (((deltaNeg * phiNegNeg - deltaNegNeg * phiNeg) / (deltaNegNeg - deltaNeg) + phi0)) / (deltaNegNeg * deltaNeg);
    const double diffPosPos =
        (((deltaPos * phiPosPos - deltaPosPos * phiPos) / (deltaPosPos - deltaPos) + phi0)) / (deltaPosPos * deltaPos);
    if (signbit(diff00) == signbit(diffPosPos)) {
      if (fabs(diffPosPos * deltaPos) < fabs(diff00 * deltaNeg)) diffPos -= deltaPos * diffPosPos;
      else diffPos += deltaNeg * diff00;
    }
    if (signbit(diff00) == signbit(diffNegNeg)) {
      if (fabs(diffNegNeg * deltaNeg) < fabs(diff00 * deltaPos)) diffNeg -= deltaNeg * diffNegNeg;
      else diffNeg += deltaPos * diff00;
    }
  }
}

// Axis values at offsets -2..2 for the first/second-order schemes (the
// second-order layer only when order == 2).
template <int D> LSG_HD inline void gatherUpwind(const BandView &b, const int32_t c[3], int order, double st[][5]) {
  if (order == 2) {
    gatherAxes<D, 2>(b, c, st);
  } else {
    double s1[D][3];
    gatherAxes<D, 1>(b, c, s1);
    LSG_UNROLL
    for (int d = 0; d < D; ++d) {
      st[d][1] = s1[d][0];
      st[d][3] = s1[d][2];
      st[d][0] = st[d][4] = 0.;
    }
  }
}

// FiniteDifferences::weno3 (lsFiniteDifferences.hpp:58-83)
LSG_HD inline double weno3(const double *x, double delta, bool plus) {
  const double eps = 1e-6;
  double d0, d1, d2;
  if (plus) {
    d0 = x[2] - x[1];
    d1 = x[3] - x[2];
    d2 = x[4] - x[3];
    double N = eps + pow2(d2 - d1);
    double Dn = eps + pow2(d1 - d0);
    double D2 = Dn * Dn;
    double wp = D2 / (D2 + 2 * N * N);
    return (d0 + d1 - wp * (d2 - 2 * d1 + d0)) / (2 * delta);
  }
  d0 = x[1] - x[0];
  d1 = x[2] - x[1];
  d2 = x[3] - x[2];
  double N = eps + pow2(d1 - d0);
  double Dn = eps + pow2(d2 - d1);
  double D2 = Dn * Dn;
  double wp = D2 / (D2 + 2 * N * N);
  return (d1 + d2 - wp * (d0 - 2 * d1 + d2)) / (2 * delta);
}

// FiniteDifferences::weno5 (lsFiniteDifferences.hpp:88-123)
LSG_HD inline double weno5(const double *x, double dx, bool plus) {
  const double eps = 1e-6;
  double d[5];
  if (!plus) {
    LSG_UNROLL
    for (int i = 0; i < 5; ++i) d[i] = x[i + 1] - x[i];
  } else {
    LSG_UNROLL
    for (int i = 0; i < 5; ++i) d[i] = x[6 - i] - x[5 - i];
  }
  double s1 = (13.0 / 12.0) * pow2(d[0] - 2 * d[1] + d[2]) + (1.0 / 4.0) * pow2(d[0] - 4 * d[1] + 3 * d[2]);
  double s2 = (13.0 / 12.0) * pow2(d[1] - 2 * d[2] + d[3]) + (1.0 / 4.0) * pow2(d[1] - d[3]);
  double s3 = (13.0 / 12.0) * pow2(d[2] - 2 * d[3] + d[4]) + (1.0 / 4.0) * pow2(3 * d[2] - 4 * d[3] + d[4]);
  double eps_scaled = eps * dx * dx;
  double al1 = 0.1 / (eps_scaled + s1);
  double al2 = 0.6 / (eps_scaled + s2);
  double al3 = 0.3 / (eps_scaled + s3);
  double alsum = al1 + al2 + al3;
  double p1 = (1.0 / 6.0) * (2 * d[0] - 7 * d[1] + 11 * d[2]);
  double p2 = (1.0 / 6.0) * (-d[1] + 5 * d[2] + 2 * d[3]);
  double p3 = (1.0 / 6.0) * (2 * d[2] + 5 * d[3] - d[4]);
  return (al1 * p1 + al2 * p2 + al3 * p3) / (alsum * dx);
}

// ------------------------------------------------- learned weno5 weights ----
// Extension of weno5: where the regularised smoothness indicators
// (eps dx^2 + beta_k) agree within a factor MLP_KINK_RATIO, the weights of
// the three candidate polynomials come from a small network fed with the
// differences d / max|d|; elsewhere (kinks) plain weno5 is used.  P holds
// the parameters as floats (bf16 values); fp32 accumulation, bias first.
// exp(x) and a / b of the network's activations: the hardware's
// single-instruction forms on the GPU (v_exp_f32, v_rcp_f32), the library
// functions in a host build.
LSG_HD inline float mlpExp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __expf(x);
#else
  return std::exp(x);
#endif
}
LSG_HD inline float mlpDiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __fdividef(a, b);
#else
  return a / b;
#endif
}
// GELU, tanh approximation 0.5 x (1 + tanh(u)) = x sigmoid(2u) with
// u = sqrt(2/pi) (x + 0.044715 x^3); sigmoid(2u) = 1 / (1 + e^-2u).
LSG_HD inline float mlpSigmoid2(float u) { return mlpDiv(1.0f, 1.0f + mlpExp(-2.0f * u)); }
LSG_HD inline float mlpGelu(float x) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float u = c * (x + a * (x * x) * x);
  return x * mlpSigmoid2(u);
}

// The forward pass of the training kernel and of the matrix-core inference:
// the inputs and both hidden activations enter their products as bf16,
// fp32 accumulation; layers 1 and 3 start from their biases, layer 2 adds
// its bias after the sum; softmax.
LSG_HD inline void mlpForward(const float *P, const float in[MLP_IN], float w[MLP_OUT]) {
  float x[MLP_IN], h1[MLP_H];
  LSG_UNROLL
  for (int i = 0; i < MLP_IN; ++i) x[i] = bf16Float(bf16Bits(in[i]));
  LSG_UNROLL
  for (int j = 0; j < MLP_H; ++j) {
    float z = P[MLP_B1 + j];
    LSG_UNROLL
    for (int i = 0; i < MLP_IN; ++i) z += P[MLP_W1 + j * MLP_IN + i] * x[i];
    h1[j] = bf16Float(bf16Bits(mlpGelu(z)));
  }
  float z3[MLP_OUT] = {P[MLP_B3], P[MLP_B3 + 1], P[MLP_B3 + 2]};
  for (int j = 0; j < MLP_H; ++j) {
    float z = 0.f;
    LSG_UNROLL
    for (int i = 0; i < MLP_H; ++i) z += P[MLP_W2 + j * MLP_H + i] * h1[i];
    const float h2 = bf16Float(bf16Bits(mlpGelu(z + P[MLP_B2 + j])));
    z3[0] += P[MLP_W3 + j] * h2;
    z3[1] += P[MLP_W3 + MLP_H + j] * h2;
    z3[2] += P[MLP_W3 + 2 * MLP_H + j] * h2;
  }
  const float m = fmaxf(z3[0], fmaxf(z3[1], z3[2]));
  const float e0 = mlpExp(z3[0] - m), e1 = mlpExp(z3[1] - m), e2 = mlpExp(z3[2] - m);
  const float sum = (e0 + e1) + e2;
  w[0] = e0 / sum;
  w[1] = e1 / sum;
  w[2] = e2 / sum;
}

LSG_HD inline double weno5Learned(const double *x, double dx, bool plus, const float *P) {
  double d[5];
  if (!plus) {
    LSG_UNROLL
    for (int i = 0; i < 5; ++i) d[i] = x[i + 1] - x[i];
  } else {
    LSG_UNROLL
    for (int i = 0; i < 5; ++i) d[i] = x[6 - i] - x[5 - i];
  }
  const double s1 = (13.0 / 12.0) * pow2(d[0] - 2 * d[1] + d[2]) + (1.0 / 4.0) * pow2(d[0] - 4 * d[1] + 3 * d[2]);
  const double s2 = (13.0 / 12.0) * pow2(d[1] - 2 * d[2] + d[3]) + (1.0 / 4.0) * pow2(d[1] - d[3]);
  const double s3 = (13.0 / 12.0) * pow2(d[2] - 2 * d[3] + d[4]) + (1.0 / 4.0) * pow2(3 * d[2] - 4 * d[3] + d[4]);
  const double e = 1e-6 * dx * dx;
  const double a1 = e + s1, a2 = e + s2, a3 = e + s3;
  const double mx = dmaxs(a1, dmaxs(a2, a3));
  const double mn = dmins(a1, dmins(a2, a3));
  double m = 0;
  LSG_UNROLL
  for (int i = 0; i < 5; ++i) m = dmaxs(m, fabs(d[i]));
  if (m == 0. || mx > MLP_KINK_RATIO * mn) return weno5(x, dx, plus);
  float in[MLP_IN];
  LSG_UNROLL
  for (int i = 0; i < 5; ++i) in[i] = (float)(d[i] / m);
  float w[MLP_OUT];
  mlpForward(P, in, w);
  const double p1 = (1.0 / 6.0) * (2 * d[0] - 7 * d[1] + 11 * d[2]);
  const double p2 = (1.0 / 6.0) * (-d[1] + 5 * d[2] + 2 * d[3]);
  const double p3 = (1.0 / 6.0) * (2 * d[2] + 5 * d[3] - d[4]);
  return ((double)w[0] * p1 + (double)w[1] * p2 + (double)w[2] * p3) / dx;
}

struct RateOut {
  double H, Dn;
};

// EngquistOsher::operator() (lsEngquistOsher.hpp:42-162)
template <int D>
LSG_HD inline RateOut schemeEO(const BandView &b, const int32_t c[3], double phi0, int order, bool calcNormals,
                               const Vel &vel) {
  double gradPos[D], gradNeg[D];
  double gradPosTotal = 0, gradNegTotal = 0;
  double st[D][5];
  gatherUpwind<D>(b, c, order, st);
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double diffPos, diffNeg;
    upwindDiffs(st[i], b.g.dx, order, phi0, diffPos, diffNeg);
    gradPos[i] = diffNeg;
    gradNeg[i] = diffPos;
    gradPosTotal += pow2(dmaxs(diffNeg, 0.0)) + pow2(dmins(diffPos, 0.0));
    gradNegTotal += pow2(dmins(diffNeg, 0.0)) + pow2(dmaxs(diffPos, 0.0));
  }
  double vel_grad = 0.;
  double n[3] = {0, 0, 0};
  if (calcNormals) {  // no zero guard (lsEngquistOsher.hpp:124-138)
    double denominator = 0;
    LSG_UNROLL
    for (int i = 0; i < D; i++) {
      double pos = st[i][3] - phi0;
      double neg = phi0 - st[i][1];
      n[i] = (pos + neg) * 0.5;
      denominator += n[i] * n[i];
    }
    denominator = 1. / sqrt(denominator);
    LSG_UNROLL
    for (int i = 0; i < D; ++i) n[i] *= denominator;
  }
  double sv = vel.scalar(n);
  if (sv > 0) vel_grad += sqrt(gradPosTotal) * sv;
  else vel_grad += sqrt(gradNegTotal) * sv;
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) vel_grad += vel.vec[w] * gradPos[w];
    else vel_grad += vel.vec[w] * gradNeg[w];
  }
  return {vel_grad, 0.};
}

// WENO::operator() (lsWENO.hpp:53-161)
// The 2D one-sided weno5 derivatives of a point with learned weights
// (weno5Learned for each): the stencil work is unrolled, the network runs in
// one loop over the evaluations that need it (one copy of its code), its
// inputs and outputs staged in the caller's scratch `io` (2D*5 floats).
// phase 1 only writes the network inputs (zeros where the network is not
// used) for a caller that runs the network for a whole wave; phase 2 then
// takes the weights from io[5e .. 5e+2] instead of running it.
template <int D>
LSG_HD inline void weno5LearnedAxes(double sts[][7], double dx, const float *P, float *io, double *gm, double *gp,
                                    int phase = 0) {
  double res[2 * D];
  double pe[2 * D][3];
  uint32_t net = 0;
  LSG_UNROLL
  for (int e = 0; e < 2 * D; ++e) {
    const double *x = sts[e >> 1];
    const bool plus = (e & 1) != 0;
    double d[5];
    if (!plus) {
      LSG_UNROLL
      for (int i = 0; i < 5; ++i) d[i] = x[i + 1] - x[i];
    } else {
      LSG_UNROLL
      for (int i = 0; i < 5; ++i) d[i] = x[6 - i] - x[5 - i];
    }
    const double s1 = (13.0 / 12.0) * pow2(d[0] - 2 * d[1] + d[2]) + (1.0 / 4.0) * pow2(d[0] - 4 * d[1] + 3 * d[2]);
    const double s2 = (13.0 / 12.0) * pow2(d[1] - 2 * d[2] + d[3]) + (1.0 / 4.0) * pow2(d[1] - d[3]);
    const double s3 = (13.0 / 12.0) * pow2(d[2] - 2 * d[3] + d[4]) + (1.0 / 4.0) * pow2(3 * d[2] - 4 * d[3] + d[4]);
    const double ee = 1e-6 * dx * dx;
    const double a1 = ee + s1, a2 = ee + s2, a3 = ee + s3;
    const double mx = dmaxs(a1, dmaxs(a2, a3));
    const double mn = dmins(a1, dmins(a2, a3));
    double m = 0;
    LSG_UNROLL
    for (int i = 0; i < 5; ++i) m = dmaxs(m, fabs(d[i]));
    res[e] = 0.;
    if (m == 0. || mx > MLP_KINK_RATIO * mn) {
      if (phase != 1) {
        res[e] = weno5(x, dx, plus);
      } else {
        LSG_UNROLL
        for (int i = 0; i < 5; ++i) io[e * 5 + i] = 0.f;
      }
    } else {
      net |= 1u << e;
      if (phase != 2) {
        LSG_UNROLL
        for (int i = 0; i < 5; ++i) io[e * 5 + i] = (float)(d[i] / m);
      }
    }
    pe[e][0] = (1.0 / 6.0) * (2 * d[0] - 7 * d[1] + 11 * d[2]);
    pe[e][1] = (1.0 / 6.0) * (-d[1] + 5 * d[2] + 2 * d[3]);
    pe[e][2] = (1.0 / 6.0) * (2 * d[2] + 5 * d[3] - d[4]);
  }
  if (phase == 1) return;
#if defined(__clang__)
#pragma unroll 1
#endif
  for (int e = 0; e < 2 * D; ++e) {
    if (phase == 2 || !((net >> e) & 1u)) continue;
    float in[MLP_IN], w[MLP_OUT];
    LSG_UNROLL
    for (int i = 0; i < MLP_IN; ++i) in[i] = io[e * 5 + i];
    mlpForward(P, in, w);
    LSG_UNROLL
    for (int k = 0; k < MLP_OUT; ++k) io[e * 5 + k] = w[k];
  }
  LSG_UNROLL
  for (int e = 0; e < 2 * D; ++e) {
    if ((net >> e) & 1u)
      res[e] = ((double)io[e * 5] * pe[e][0] + (double)io[e * 5 + 1] * pe[e][1] + (double)io[e * 5 + 2] * pe[e][2]) / dx;
    if (e & 1) gp[e >> 1] = res[e];
    else gm[e >> 1] = res[e];
  }
}

// Learned weno5 in two phases can keep the axis values of the first in
// `stc` (element q of the D x 7 values at stc[q * stcStride]): the second
// phase then reads them there instead of gathering them again.
template <int D, int ORDER, bool LEARNED = false>
LSG_HD inline RateOut schemeWENO(const BandView &b, const int32_t c[3], double phi0, bool calcNormals, const Vel &vel,
                                 const float *mlp = nullptr, float *io = nullptr, int phase = 0,
                                 const double *stc = nullptr, int stcStride = 0) {
  constexpr int r = (ORDER + 1) / 2;
  double gradPosTotal = 0, gradNegTotal = 0;
  double gm[D], gp[D];
  double sts[D][2 * r + 1];
  if (LEARNED && phase == 2 && stc) {
    LSG_UNROLL
    for (int d = 0; d < D; ++d)
      LSG_UNROLL
      for (int k = 0; k < 2 * r + 1; ++k) sts[d][k] = stc[(d * (2 * r + 1) + k) * stcStride];
  } else {
    gatherAxes<D, r>(b, c, sts);
  }
  double cd[D];  // central differences of the normal (lsWENO.hpp:113-129)
  LSG_UNROLL
  for (int i = 0; i < D; i++) cd[i] = sts[i][r + 1] - sts[i][r - 1];
  if constexpr (LEARNED) {
    LSG_UNROLL
    for (int i = 0; i < D; i++) sts[i][r] = phi0;
    weno5LearnedAxes<D>(sts, b.g.dx, mlp, io, gm, gp, phase);
    if (phase == 1) return {0., 0.};
  }
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double *st = sts[i];
    st[r] = phi0;
    if (ORDER == 3) {
      gm[i] = weno3(st, b.g.dx, false);
      gp[i] = weno3(st, b.g.dx, true);
    } else if (!LEARNED) {
      gm[i] = weno5(st, b.g.dx, false);
      gp[i] = weno5(st, b.g.dx, true);
    }
    gradPosTotal += pow2(dmaxs(gm[i], 0.0)) + pow2(dmins(gp[i], 0.0));
    gradNegTotal += pow2(dmins(gm[i], 0.0)) + pow2(dmaxs(gp[i], 0.0));
  }
  double vel_grad = 0.;
  double n[3] = {0, 0, 0};
  if (calcNormals) {  // central difference with zero guard (lsWENO.hpp:113-129)
    double denominator = 0;
    LSG_UNROLL
    for (int i = 0; i < D; i++) {
      n[i] = cd[i];
      denominator += n[i] * n[i];
    }
    if (denominator > 0) {
      denominator = 1. / sqrt(denominator);
      LSG_UNROLL
      for (int i = 0; i < D; ++i) n[i] *= denominator;
    }
  }
  double sv = vel.scalar(n);
  if (sv > 0) vel_grad += sqrt(gradPosTotal) * sv;
  else vel_grad += sqrt(gradNegTotal) * sv;
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) vel_grad += vel.vec[w] * gm[w];
    else vel_grad += vel.vec[w] * gp[w];
  }
  return {vel_grad, 0.};
}

// LaxFriedrichs::operator() (lsLaxFriedrichs.hpp:46-158)
template <int D>
LSG_HD inline RateOut schemeLF(const BandView &b, const int32_t c[3], double phi0, int order, bool calcNormals,
                               const Vel &vel, double alphaFactor, const double *alphas) {
  double gradPos[D], gradNeg[D];
  double grad = 0., dissipation = 0.;
  double n[3] = {0, 0, 0};
  double normalModulus = 0;
  double st[D][5];
  gatherUpwind<D>(b, c, order, st);
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double diffPos, diffNeg;
    upwindDiffs(st[i], b.g.dx, order, phi0, diffPos, diffNeg);
    gradPos[i] = diffNeg;
    gradNeg[i] = diffPos;
    if (calcNormals) {
      n[i] = (diffNeg + diffPos) * 0.5;
      normalModulus += n[i] * n[i];
    }
    grad += pow2((diffNeg + diffPos) * 0.5);
    dissipation += alphaFactor * alphas[i] * (diffPos - diffNeg) * 0.5;
  }
  if (calcNormals) {
    normalModulus = 1. / sqrt(normalModulus);
    LSG_UNROLL
    for (int i = 0; i < D; ++i) n[i] *= normalModulus;
  }
  double sv = vel.scalar(n);
  double totalGrad = 0.;
  if (sv != 0.) totalGrad = sv * sqrt(grad);
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) totalGrad += vel.vec[w] * gradPos[w];
    else totalGrad += vel.vec[w] * gradNeg[w];
  }
  return {totalGrad, (totalGrad != 0.) ? dissipation : 0};
}

// LocalLocalLaxFriedrichs::operator() (lsLocalLocalLaxFriedrichs.hpp:42-165)
template <int D>
LSG_HD inline RateOut schemeLLLF(const BandView &b, const int32_t c[3], double phi0, int order, const Vel &vel,
                                 double alphaFactor, double *alphaAcc) {
  double gradPos[D], gradNeg[D];
  double grad = 0., dissipation = 0.;
  double n[3] = {0, 0, 0};
  double normalModulus = 0;
  double st[D][5];
  gatherUpwind<D>(b, c, order, st);
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double diffPos, diffNeg;
    upwindDiffs(st[i], b.g.dx, order, phi0, diffPos, diffNeg);
    gradPos[i] = diffNeg;
    gradNeg[i] = diffPos;
    n[i] = (diffNeg + diffPos) * 0.5;
    normalModulus += n[i] * n[i];
    grad += pow2((diffNeg + diffPos) * 0.5);
  }
  normalModulus = 1. / sqrt(normalModulus);
  LSG_UNROLL
  for (int i = 0; i < D; ++i) n[i] *= normalModulus;
  double sv = vel.scalar(n);
  double totalGrad = 0.;
  if (sv != 0.) totalGrad = sv * sqrt(grad);
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) totalGrad += vel.vec[w] * gradPos[w];
    else totalGrad += vel.vec[w] * gradNeg[w];
  }
  if (totalGrad == 0.) return {0., 0.};
  LSG_UNROLL
  for (int i = 0; i < D; ++i) {
    double alpha = fabs((sv + vel.vec[i]) * n[i]);
    alphaAcc[i] = dmaxs(alphaAcc[i], alpha);
    dissipation += alphaFactor * alpha * (gradNeg[i] - gradPos[i]) * 0.5;
  }
  return {totalGrad, dissipation};
}

// LocalLaxFriedrichsAnalytical<T,D,1>::operator()
// (lsLocalLaxFriedrichsAnalytical.hpp:66-229): LLLF's Hamiltonian with the
// velocity field's analytical alphas (the built-in functor's are constant
// per direction, so the 3^D box of normals does not enter), no dissipation
// factor.
template <int D>
LSG_HD inline RateOut schemeLLFA(const BandView &b, const int32_t c[3], double phi0, const Vel &vel, double *alphaAcc) {
  double gradPos[D], gradNeg[D];
  double grad = 0., dissipation = 0.;
  double n[3] = {0, 0, 0};
  double normalModulus = 0;
  double st[D][5];
  gatherUpwind<D>(b, c, 1, st);
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double diffPos, diffNeg;
    upwindDiffs(st[i], b.g.dx, 1, phi0, diffPos, diffNeg);
    gradPos[i] = diffNeg;
    gradNeg[i] = diffPos;
    n[i] = (diffNeg + diffPos) * 0.5;
    normalModulus += n[i] * n[i];
    grad += pow2((diffNeg + diffPos) * 0.5);
  }
  normalModulus = 1. / sqrt(normalModulus);
  LSG_UNROLL
  for (int i = 0; i < D; ++i) n[i] *= normalModulus;
  double sv = vel.scalar(n);
  double totalGrad = 0.;
  if (sv != 0.) totalGrad = sv * sqrt(grad);
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) totalGrad += vel.vec[w] * gradPos[w];
    else totalGrad += vel.vec[w] * gradNeg[w];
  }
  if (totalGrad == 0.) return {0., 0.};
  LSG_UNROLL
  for (int i = 0; i < D; ++i) {
    // the maximum of getDissipationAlpha over the 3^D box (constant for the
    // built-in functors; the host's values at the box points otherwise)
    double a = vel.alpha[i];
    if (vel.box) {
      constexpr int NB = (D == 3) ? 27 : 9;
      a = vel.box[1 + i];
      for (int q = 1; q < NB; ++q) a = dmaxs(a, vel.box[q * vel.boxStride + 1 + i]);
    }
    const double alpha = dmaxs(0.0, a);
    alphaAcc[i] = dmaxs(alphaAcc[i], a);
    dissipation += alpha * (gradNeg[i] - gradPos[i]) * 0.5;
  }
  return {totalGrad, dissipation};
}

// The central differences of the analytical local Lax-Friedrichs alpha box
// at box point q around c (lsLocalLaxFriedrichsAnalytical.hpp:191-208,
// calculateNormalComponent :30-34): ((pos - centre) / dx + (centre - neg)
// / dx) / 2 per direction, the neighbours boundary mapped per coordinate of
// their offset from c, not normalised.
template <int D>
LSG_HD inline void llfaBoxDiffs(const BandView &b, const int32_t c[3], int q, double out[3]) {
  const int ni[3] = {q % 3 - 1, (q / 3) % 3 - 1, q / 9 - 1};
  BandView::Loc l[2 * D + 1];
  int32_t nc[3] = {0, 0, 0};
  LSG_UNROLL
  for (int d = 0; d < D; ++d) nc[d] = b.g.map(d, (int64_t)c[d] + ni[d]);
  l[2 * D] = b.locate(nc[0], nc[1], nc[2]);
  LSG_UNROLL
  for (int dir = 0; dir < D; ++dir) {
    int32_t pp[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
    LSG_UNROLL
    for (int d = 0; d < D; ++d) {
      pp[d] = b.g.map(d, (int64_t)c[d] + ni[d] + (d == dir ? 1 : 0));
      pn[d] = b.g.map(d, (int64_t)c[d] + ni[d] - (d == dir ? 1 : 0));
    }
    l[dir] = b.locate(pp[0], pp[1], pp[2]);
    l[D + dir] = b.locate(pn[0], pn[1], pn[2]);
  }
  double v[2 * D + 1];
  b.values<2 * D + 1>(l, v);
  out[0] = out[1] = out[2] = 0.;
  LSG_UNROLL
  for (int dir = 0; dir < D; ++dir) {
    const double diffPos = (v[dir] - v[2 * D]) / b.g.dx;
    const double diffNeg = (v[2 * D] - v[D + dir]) / b.g.dx;
    out[dir] = (diffPos + diffNeg) * 0.5;
  }
}

// Normal at point q of the local Lax-Friedrichs alpha box around c
// (lsLocalLaxFriedrichs.hpp:193-215): central differences of the boundary
// mapped neighbours, normalised; q = (ox + 1) + 3 (oy + 1) + 9 (oz + 1).
// The 2D + 1 lookups of box point q's normal: slot 2D the box point, dir
// and D + dir its neighbours along dir (+, -).  With lin >= 0 (mappedLin
// with R = 2) every point is a frame bit at its mapped offset; otherwise
// each is located.
template <int D>
LSG_HD inline void llfBoxLocs(const BandView &b, const int32_t c[3], int q, int64_t lin, const int64_t off[][5],
                              BandView::Loc *l) {
  const int ni[3] = {q % 3 - 1, (q / 3) % 3 - 1, q / 9 - 1};
  if (lin >= 0) {
    int64_t base = lin;
    LSG_UNROLL
    for (int d = 0; d < D; ++d) base += off[d][2 + ni[d]];
    l[2 * D] = locAt(base);
    LSG_UNROLL
    for (int dir = 0; dir < D; ++dir) {
      const int64_t here = off[dir][2 + ni[dir]];
      l[dir] = locAt(base - here + off[dir][3 + ni[dir]]);
      l[D + dir] = locAt(base - here + off[dir][1 + ni[dir]]);
    }
    return;
  }
  int32_t nc[3] = {0, 0, 0};
  LSG_UNROLL
  for (int d = 0; d < D; ++d) nc[d] = b.g.map(d, (int64_t)c[d] + ni[d]);
  l[2 * D] = b.locate(nc[0], nc[1], nc[2]);
  LSG_UNROLL
  for (int dir = 0; dir < D; ++dir) {
    int32_t pp[3] = {0, 0, 0}, pn[3] = {0, 0, 0};
    LSG_UNROLL
    for (int d = 0; d < D; ++d) {
      pp[d] = b.g.map(d, (int64_t)c[d] + ni[d] + (d == dir ? 1 : 0));
      pn[d] = b.g.map(d, (int64_t)c[d] + ni[d] - (d == dir ? 1 : 0));
    }
    l[dir] = b.locate(pp[0], pp[1], pp[2]);
    l[D + dir] = b.locate(pn[0], pn[1], pn[2]);
  }
}
// The normal from those values: central differences, normalised.
template <int D> LSG_HD inline void llfBoxNormalFrom(const double *bv, double dx, double normal[3]) {
  normal[0] = normal[1] = normal[2] = 0.;
  double nm = 0.;
  const double center = bv[2 * D];
  LSG_UNROLL
  for (int dir = 0; dir < D; ++dir) {
    const double neg = bv[D + dir];
    const double pos = bv[dir];
    const double dP = (pos - center) / dx;
    const double dN = (center - neg) / dx;
    normal[dir] = (dP + dN) * 0.5;
    nm += normal[dir] * normal[dir];
  }
  nm = 1. / sqrt(nm);
  LSG_UNROLL
  for (int dir = 0; dir < D; ++dir) normal[dir] *= nm;
}
template <int D> LSG_HD inline void llfBoxNormal(const BandView &b, const int32_t c[3], int q, double normal[3]) {
  BandView::Loc l[2 * D + 1];
  llfBoxLocs<D>(b, c, q, -1, nullptr, l);
  double bv[2 * D + 1];
  b.values<2 * D + 1>(l, bv);
  llfBoxNormalFrom<D>(bv, b.g.dx, normal);
}

// LocalLaxFriedrichs::operator() (lsLocalLaxFriedrichs.hpp:65-241)
template <int D>
LSG_HD inline RateOut schemeLLF(const BandView &b, const int32_t c[3], double phi0, int order, const Vel &vel,
                                double alphaFactor, double *alphaAcc) {
  double gradPos[D], gradNeg[D];
  double grad = 0., dissipation = 0.;
  double n[3] = {0, 0, 0};
  double normalModulus = 0;
  double st[D][5];
  gatherUpwind<D>(b, c, order, st);
  LSG_UNROLL
  for (int i = 0; i < D; i++) {
    double diffPos, diffNeg;
    upwindDiffs(st[i], b.g.dx, order, phi0, diffPos, diffNeg);
    gradPos[i] = diffNeg;
    gradNeg[i] = diffPos;
    n[i] = (diffNeg + diffPos) * 0.5;
    normalModulus += n[i] * n[i];
    grad += pow2((diffNeg + diffPos) * 0.5);
  }
  normalModulus = 1. / sqrt(normalModulus);
  LSG_UNROLL
  for (int i = 0; i < D; ++i) n[i] *= normalModulus;
  double sv = vel.scalar(n);
  double totalGrad = 0.;
  if (sv != 0.) totalGrad = sv * sqrt(grad);
  LSG_UNROLL
  for (int w = 0; w < D; w++) {
    if (vel.vec[w] > 0.) totalGrad += vel.vec[w] * gradPos[w];
    else totalGrad += vel.vec[w] * gradNeg[w];
  }
  if (totalGrad == 0.) return {0., 0.};
  double alpha[D];
  LSG_UNROLL
  for (int i = 0; i < D; ++i) alpha[i] = 0;
  constexpr int numNeighbors = (D == 3) ? 27 : 9;
  // box loop (lsLocalLaxFriedrichs.hpp:193-231), x fastest; a box point's
  // 2D + 1 lookups are at fixed offsets from c (mappedLin, radius 2) away
  // from the frame edge.  G box points per batch of lookups: 1 measured
  // best (rough1024 LLF2: 534 us per rate pass against 742 us with 3, whose
  // registers halve the occupancy)
  int64_t off[3][5];
  const int64_t lin = mappedLin<D, 2>(b, c, off);
  constexpr int G = 1, NL = 2 * D + 1;
#if defined(__clang__)
#pragma unroll 1
#endif
  for (int q0 = 0; q0 < numNeighbors; q0 += G) {
    BandView::Loc l[G * NL];
    LSG_UNROLL
    for (int g = 0; g < G; ++g) llfBoxLocs<D>(b, c, q0 + g, lin, off, l + g * NL);
    double bv[G * NL];
    b.values<G * NL>(l, bv);
    LSG_UNROLL
    for (int g = 0; g < G; ++g) {
      const int q = q0 + g;
      double normal[3];
      llfBoxNormalFrom<D>(bv + g * NL, b.g.dx, normal);
      // the box velocities: the functor at the box normal, or the host's
      // values for this box point
      double sv, vv[3];
      if (vel.box) {
        const double *h = vel.box + q * vel.boxStride;
        sv = h[0];
        vv[0] = h[1];
        vv[1] = h[2];
        vv[2] = h[3];
      } else {
        sv = vel.scalar(normal);
        vv[0] = vel.vec[0];
        vv[1] = vel.vec[1];
        vv[2] = vel.vec[2];
      }
      LSG_UNROLL
      for (int dir = 0; dir < D; ++dir) {
        double tempAlpha = fabs((sv + vv[dir]) * normal[dir]);
        alpha[dir] = dmaxs(alpha[dir], tempAlpha);
        alphaAcc[dir] = dmaxs(alphaAcc[dir], tempAlpha);
      }
    }
  }
  LSG_UNROLL
  for (int i = 0; i < D; ++i) dissipation += alphaFactor * alpha[i] * (gradNeg[i] - gradPos[i]) * 0.5;
  return {totalGrad, dissipation};
}

// Values of the 3^D box around c, boundary mapped per coordinate of the
// offset from the centre: box[(ox + 1) + 3 (oy + 1) + 9 (oz + 1)].  All
// lookups are issued together.
template <int D> LSG_HD inline void gatherBox(const BandView &b, const int32_t c[3], double *box) {
  constexpr int NB = (D == 3) ? 27 : 9;
  BandView::Loc l[NB];
  int64_t off[3][3];
  const int64_t lin = mappedLin<D, 1>(b, c, off);
  if (lin >= 0) {
    LSG_UNROLL
    for (int q = 0; q < NB; ++q)
      l[q] = locAt(lin + off[0][q % 3] + off[1][(q / 3) % 3] + (D == 3 ? off[2][q / 9] : 0));
  } else {
    LSG_UNROLL
    for (int q = 0; q < NB; ++q) {
      const int o[3] = {q % 3 - 1, (q / 3) % 3 - 1, q / 9 - 1};
      int32_t p[3] = {0, 0, 0};
      LSG_UNROLL
      for (int d = 0; d < D; ++d) p[d] = b.g.map(d, (int64_t)c[d] + o[d]);
      l[q] = b.locate(p[0], p[1], p[2]);
    }
  }
  b.values<NB>(l, box);
}

// StencilLocalLaxFriedrichsScalar<T,D,1>::operator()
// (lsStencilLocalLaxFriedrichsScalar.hpp:238-383), first-order differences
// and the simple adaptive epsilon (:62-74).  The normal and gradient at a
// box offset differentiate along axis i over the row through the offset with
// coordinate i at -1, 0, +1 (:142-186), so the 3^D box is the whole support.
// A box point whose normal is below 1e-6 in every component ends the alpha
// loop (its `continue` skips incrementIndex: the later slots stay zero).
// `used` is set when the point sets finalAlphas (scalar velocity != 0).
#define LSG_BOX(x, y, z) box[((x) + 1) + 3 * ((y) + 1) + (D == 3 ? 9 * ((z) + 1) : 0)]
// FiniteDifferences<T, FIRST_ORDER>::calculateGradient at a box offset
// (lsFiniteDifferences.hpp:177-181), over the 3^D box values of gatherBox
template <int D>
LSG_HD inline void sllfGradient(const double *box, double dx, int ox, int oy, int oz, double *g) {
  g[0] = ((LSG_BOX(1, oy, oz) - LSG_BOX(0, oy, oz)) / dx + (LSG_BOX(0, oy, oz) - LSG_BOX(-1, oy, oz)) / dx) * 0.5;
  g[1] = ((LSG_BOX(ox, 1, oz) - LSG_BOX(ox, 0, oz)) / dx + (LSG_BOX(ox, 0, oz) - LSG_BOX(ox, -1, oz)) / dx) * 0.5;
  g[2] = 0.;
  if (D == 3)
    g[2] = ((LSG_BOX(ox, oy, 1) - LSG_BOX(ox, oy, 0)) / dx + (LSG_BOX(ox, oy, 0) - LSG_BOX(ox, oy, -1)) / dx) * 0.5;
}
// calculateNormal (lsStencilLocalLaxFriedrichsScalar.hpp:142-163)
template <int D>
LSG_HD inline void sllfNormal(const double *box, double dx, int ox, int oy, int oz, double *n) {
  double g[3];
  sllfGradient<D>(box, dx, ox, oy, oz, g);
  double modulus = 0.;
  LSG_UNROLL
  for (int i = 0; i < D; ++i) modulus += g[i] * g[i];
  modulus = 1. / sqrt(modulus);
  n[0] = n[1] = n[2] = 0.;
  LSG_UNROLL
  for (int i = 0; i < D; ++i) n[i] = g[i] * modulus;
}
// A box normal below 1e-6 in every component (:283-288)
LSG_HD inline bool sllfCorrupt(const double n[3]) {
  return fabs(n[0]) < 1e-6 && fabs(n[1]) < 1e-6 && fabs(n[2]) < 1e-6;
}
// Normal offset of the simple adaptive epsilon (:62-74) for a box point's
// scalar velocity
LSG_HD inline double sllfEpsilon(double baseEps, double lsv) {
  const double absVel = fabs(lsv);
  double DN = baseEps;
  if (absVel > 1e-10) DN = baseEps * dmaxs(0.1, dmins(10.0, absVel));
  return DN;
}

template <int D>
LSG_HD inline RateOut schemeSLLF(const BandView &b, const int32_t c[3], const Vel &vel, double baseNormalEps,
                                 double *finalAlphas, bool &used) {
  constexpr int NB = (D == 3) ? 27 : 9;
  double box[NB];
  gatherBox<D>(b, c, box);
  const double dx = b.g.dx;
  used = false;
  auto gradient = [&](int ox, int oy, int oz, double *g) { sllfGradient<D>(box, dx, ox, oy, oz, g); };
  auto normalAt = [&](int ox, int oy, int oz, double *n) { sllfNormal<D>(box, dx, ox, oy, oz, n); };
  double normal[3];
  normalAt(0, 0, 0, normal);
  double sv = vel.scalar(normal);
  LSG_UNROLL
  for (int i = 0; i < D; ++i) sv += vel.vec[i] * normal[i];
  if (sv == 0.) return {0., 0.};
  double g0[3];
  gradient(0, 0, 0, g0);
  double dot = 0;
  LSG_UNROLL
  for (int i = 0; i < D; ++i) dot += g0[i] * g0[i];
  const double hamiltonian = sqrt(dot) * sv;
  double maxAlpha[3] = {0., 0., 0.};
  const double baseEps = baseNormalEps * dx;
  bool alive = true;  // cleared by the first box point with a corrupted normal
  LSG_UNROLL
  for (int q = 0; q < NB; ++q) {  // incrementIndex order (:212-220): x fastest
    const int ox = q % 3 - 1, oy = (q / 3) % 3 - 1, oz = (D == 3) ? q / 9 - 1 : 0;
    double ln[3];
    normalAt(ox, oy, oz, ln);
    alive = alive && !sllfCorrupt(ln);
    if (!alive) continue;
    // the box point's velocities: the functor's, or the host's box record
    const double *hb = vel.box ? vel.box + q * vel.boxStride : nullptr;
    double lsv = hb ? hb[0] : vel.scalar(ln);
    LSG_UNROLL
    for (int d = 0; d < D; ++d) lsv += (hb ? hb[1 + d] : vel.vec[d]) * ln[d];
    const double DN = sllfEpsilon(baseEps, lsv);
    double np[3] = {ln[0], ln[1], ln[2]}, nn[3] = {ln[0], ln[1], ln[2]};
    double vd[3] = {0., 0., 0.};
    LSG_UNROLL
    for (int k = 0; k < D; ++k) {  // :324-338
      np[k] -= DN;
      nn[k] += DN;
      const double vp = vel.sllf ? vel.sllf[(q * D + k) * 2] : vel.scalar(np);
      const double vn = vel.sllf ? vel.sllf[(q * D + k) * 2 + 1] : vel.scalar(nn);
      vd[k] = (vn - vp) / (2.0 * DN);
      np[k] += DN;
      nn[k] -= DN;
    }
    double g[3];
    gradient(ox, oy, oz, g);
    double den = 0;
    LSG_UNROLL
    for (int i = 0; i < D; ++i) den += g[i] * g[i];
    LSG_UNROLL
    for (int k = 0; k < D; ++k) {  // :341-364
      double monti = 0, toifl = 0;
      LSG_UNROLL
      for (int j = 0; j < D - 1; ++j) {
        const int idx = (k + 1 + j) % D;
        monti += g[idx] * g[idx];
        toifl += g[idx] * vd[idx];
      }
      monti *= vd[k] / den;
      toifl *= -g[k] / den;
      const double osher = lsv * ln[k];
      maxAlpha[k] = dmaxs(maxAlpha[k], fabs(monti + toifl + osher));  // :371-376
    }
  }
  double dissipation = 0.;
  LSG_UNROLL
  for (int d = 0; d < D; ++d) {  // :369-380, calculateGradientDiff at the centre
    const double v0 = (d == 0) ? LSG_BOX(-1, 0, 0) : (d == 1) ? LSG_BOX(0, -1, 0) : LSG_BOX(0, 0, -1);
    const double v2 = (d == 0) ? LSG_BOX(1, 0, 0) : (d == 1) ? LSG_BOX(0, 1, 0) : LSG_BOX(0, 0, 1);
    const double v1 = LSG_BOX(0, 0, 0);
    const double gd = ((v2 - v1) / dx - (v1 - v0) / dx) * 0.5;
    finalAlphas[d] = dmins(maxAlpha[d], 1.7976931348623157e308);  // min(maxAlpha, maxDissipation)
    dissipation += finalAlphas[d] * gd;
  }
#undef LSG_BOX
  used = true;
  return {hamiltonian, dissipation};
}

// Normal vector a scheme passes to VelocityField::getScalarVelocity at an
// active point (host-evaluated velocities): EO without a zero guard
// (lsEngquistOsher.hpp:124-138), WENO with one (lsWENO.hpp:113-129), both
// only with calcNormals; LLLF always, from the order-corrected one-sided
// differences (lsLocalLocalLaxFriedrichs.hpp:95-120).
template <int D> LSG_HD inline void schemeNormal(const BandView &b, const int32_t c[3], double phi0, int scheme,
                                                 int order, bool calcNormals, double n[3]) {
  n[0] = n[1] = n[2] = 0.;
  // upwind-difference normal: LLLF, LLF and the analytical LLF (first
  // order) always (lsLocalLocalLaxFriedrichs.hpp:118-124,
  // lsLocalLaxFriedrichs.hpp:140-155, lsLocalLaxFriedrichsAnalytical.hpp:
  // 122-159), LF with calculateNormalVectors (lsLaxFriedrichs.hpp:127-141)
  if (scheme == 3 /* S_LLLF */ || scheme == 4 /* S_LLF */ || scheme == 7 /* S_LLFA */ ||
      (scheme == 2 /* S_LF */ && calcNormals)) {
    const int ord = (scheme == 7) ? 1 : order;
    double st[D][5];
    gatherUpwind<D>(b, c, ord, st);
    double nm = 0;
    LSG_UNROLL
    for (int i = 0; i < D; i++) {
      double diffPos, diffNeg;
      upwindDiffs(st[i], b.g.dx, ord, phi0, diffPos, diffNeg);
      n[i] = (diffNeg + diffPos) * 0.5;
      nm += n[i] * n[i];
    }
    nm = 1. / sqrt(nm);
    LSG_UNROLL
    for (int i = 0; i < D; ++i) n[i] *= nm;
    return;
  }
  if (scheme == 8 /* S_SLLF */) {  // calculateNormal at the centre (lsStencilLocalLaxFriedrichsScalar.hpp:252)
    double box[(D == 3) ? 27 : 9];
    gatherBox<D>(b, c, box);
    sllfNormal<D>(box, b.g.dx, 0, 0, 0, n);
    return;
  }
  if (!calcNormals || scheme == 2 /* S_LF */) return;
  double st[D][3];
  gatherAxes<D, 1>(b, c, st);
  double den = 0;
  if (scheme == 0 /* S_EO */) {
    LSG_UNROLL
    for (int i = 0; i < D; i++) {
      const double pos = st[i][2] - phi0;
      const double neg = phi0 - st[i][0];
      n[i] = (pos + neg) * 0.5;
      den += n[i] * n[i];
    }
    den = 1. / sqrt(den);
    LSG_UNROLL
    for (int i = 0; i < D; ++i) n[i] *= den;
  } else {  // WENO
    LSG_UNROLL
    for (int i = 0; i < D; i++) {
      n[i] = st[i][2] - st[i][0];
      den += n[i] * n[i];
    }
    if (den > 0) {
      den = 1. / sqrt(den);
      LSG_UNROLL
      for (int i = 0; i < D; ++i) n[i] *= den;
    }
  }
}

// Scheme ids of the rate kernel.
enum SchemeId : int {
  S_EO = 0, S_WENO5 = 1, S_LF = 2, S_LLLF = 3, S_LLF = 4, S_WENO3 = 5, S_WENO5L = 6, S_LLFA = 7, S_SLLF = 8
};

LSG_HD inline int schemeIdOf(int spatial) {
  switch (spatial) {
  case 0: case 1: return S_EO;
  case 2: case 3: return S_LF;
  case 4: return S_LLFA;
  case 5: case 6: return S_LLLF;
  case 7: case 8: return S_LLF;
  case 9: return S_SLLF;
  case 10: return S_WENO3;
  case 12: return S_WENO5L;
  default: return S_WENO5;
  }
}

// Hamiltonian and dissipation of one active point; returns the stage
// velocity after updateLevelSet's checkDissipation rule (lsAdvect.hpp:738-744,
// precedence kept) and the raw H - D through `velocity`.  SLLF writes the
// point's finalAlphas to alphaAcc and sets *sllfUsed when it set them.
template <int D, int SCHEME>
LSG_HD inline double ratePoint(const BandView &b, const int32_t c[3], double phi0, const RateParams &rp, const Vel &vel,
                               const double *globalAlphas, double *alphaAcc, double &velocity,
                               const float *mlp = nullptr, float *io = nullptr, bool *sllfUsed = nullptr,
                               double *dnOut = nullptr, int learnedPhase = 0, const double *stc = nullptr,
                               int stcStride = 0) {
  RateOut r;
  if (SCHEME == S_SLLF) {
    bool used = false;
    r = schemeSLLF<D>(b, c, vel, rp.sllfBaseEps, alphaAcc, used);
    if (sllfUsed) *sllfUsed = used;
  } else if (SCHEME == S_EO) r = schemeEO<D>(b, c, phi0, rp.order, rp.calcNormals != 0, vel);
  else if (SCHEME == S_WENO5) r = schemeWENO<D, 5>(b, c, phi0, rp.calcNormals != 0, vel);
  else if (SCHEME == S_WENO3) r = schemeWENO<D, 3>(b, c, phi0, rp.calcNormals != 0, vel);
  else if (SCHEME == S_WENO5L) r = schemeWENO<D, 5, true>(b, c, phi0, rp.calcNormals != 0, vel, mlp, io, learnedPhase,
                                                             stc, stcStride);
  else if (SCHEME == S_LF) r = schemeLF<D>(b, c, phi0, rp.order, rp.calcNormals != 0, vel, rp.alphaFactor, globalAlphas);
  else if (SCHEME == S_LLLF) r = schemeLLLF<D>(b, c, phi0, rp.order, vel, rp.alphaFactor, alphaAcc);
  else if (SCHEME == S_LLFA) r = schemeLLFA<D>(b, c, phi0, vel, alphaAcc);
  else r = schemeLLF<D>(b, c, phi0, rp.order, vel, rp.alphaFactor, alphaAcc);
  velocity = r.H - r.Dn;
  if (dnOut) *dnOut = r.Dn;
  double v = velocity;
  if ((rp.checkDissipation && (r.H < 0 && v > 0)) || (r.H > 0 && v < 0)) v = 0;
  return v;
}

// Normal of Advect::findGlobalAlphas (lsAdvect.hpp:150-162): phiPos -
// phiNeg per direction, normalised.
template <int D> LSG_HD inline void globalAlphaNormal(const BandView &b, const int32_t c[3], double normal[3]) {
  normal[0] = normal[1] = normal[2] = 0.;
  double nm = 0.;
  double st[D][3];
  gatherAxes<D, 1>(b, c, st);
  LSG_UNROLL
  for (int k = 0; k < D; ++k) {
    normal[k] = st[k][2] - st[k][0];
    nm += normal[k] * normal[k];
  }
  nm = 1. / sqrt(nm);
  LSG_UNROLL
  for (int k = 0; k < D; ++k) normal[k] *= nm;
}
// Per-point alphas of Advect::findGlobalAlphas (lsAdvect.hpp:150-178).
template <int D> LSG_HD inline void globalAlphaPoint(const BandView &b, const int32_t c[3], const Vel &vel, double *la) {
  double normal[3];
  globalAlphaNormal<D>(b, c, normal);
  const double sv = vel.scalar(normal);
  LSG_UNROLL
  for (int k = 0; k < D; ++k) {
    double t = fabs((sv + vel.vec[k]) * normal[k]);
    la[k] = dmaxs(la[k], t);
  }
}

// Time step of a rate pass from max |H - D| (min over points of cfl/|v|
// equals cfl / max|v| since correctly rounded division is monotone) and the
// reduceTimeStepHamiltonJacobi bound of the LF family (lsAdvect.hpp:530-606,
// lsLaxFriedrichs.hpp:160-172).
// Time step of a rate pass with several level sets: the minimum over the
// active points of their summed step limits (lsAdvect.hpp:591-592), then the
// LF-family bound as in stageTimeStep.
LSG_HD inline double stageTimeStepMM(double maxTimeStep, double minStep, int D, const double *alphas, bool useAlphas,
                                     double dx) {
  double tempMax = maxTimeStep;
  if (minStep < tempMax) tempMax = minStep;
  if (useAlphas) {
    double timeStep = 0;
    LSG_UNROLL
    for (int i = 0; i < D; ++i) timeStep += alphas[i] / dx;
    timeStep = 1.0 / timeStep;
    tempMax = (tempMax < timeStep) ? tempMax : timeStep;
  }
  double res = maxTimeStep;
  if (tempMax < res) res = tempMax;
  return res;
}

LSG_HD inline double stageTimeStep(double maxTimeStep, double cfl, double maxAbs, int D, const double *alphas,
                                   bool useAlphas, double dx) {
  double tempMax = maxTimeStep;
  if (maxAbs > 0.) {
    double cand = cfl / maxAbs;
    if (cand < tempMax) tempMax = cand;
  }
  if (useAlphas) {
    double timeStep = 0;
    LSG_UNROLL
    for (int i = 0; i < D; ++i) timeStep += alphas[i] / dx;
    timeStep = 1.0 / timeStep;
    tempMax = (tempMax < timeStep) ? tempMax : timeStep;
  }
  double res = maxTimeStep;
  if (tempMax < res) res = tempMax;
  return res;
}

}  // namespace lsg
