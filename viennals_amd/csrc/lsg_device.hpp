// Device-side level-set band layout for the MI355X advection path.
//
// A band is the canonical run-start list of the level-set value field (every
// defined point plus every undefined point that starts a run of a new sign,
// in HRLE order: x fastest, highest dimension most significant; see
// include/viennals/lsDomain.hpp:51-52 for the +-DBL_MAX undefined values).
// Entries inside a rectangular "frame" are indexed by a bitmap with one bit
// per grid point (rows padded to a power of two) and a per-word exclusive
// prefix count, so the rank of any grid point -- the index of the run that
// contains it -- costs one 64-bit word load, one prefix load and a popcount.
// Entry 0 is always the grid minimum point; when it lies outside the frame
// (any infinite dimension) it is kept separately and `base` is 1.
#pragma once

#ifndef LSG_HOST_ONLY
#include <hip/hip_runtime.h>
#define LSG_HD __host__ __device__
#else
#define LSG_HD
#endif

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

// Full unrolling of the short per-dimension loops keeps the small local
// arrays (stencils, neighbour values, coordinates) in registers.
#if defined(__clang__)
#define LSG_UNROLL _Pragma("unroll")
#else
#define LSG_UNROLL
#endif

namespace lsg {

#ifdef LSG_HOST_ONLY
using std::fabs;
using std::signbit;
using std::sqrt;
#endif

constexpr double POS = 1.7976931348623157e308;   // Domain::POS_VALUE
constexpr double NEG = -1.7976931348623157e308;  // Domain::NEG_VALUE
constexpr int32_t INF_IDX = 715827882;          // HRLE infinite extent (INT32_MAX/3)
constexpr double DEPS = 2.220446049250313e-16;   // numeric_limits<double>::epsilon()

enum BC : int { REFLECTIVE = 0, INFINITE = 1, PERIODIC = 2, POS_INFINITE = 3, NEG_INFINITE = 4 };

LSG_HD inline bool isUndef(double v) { return v == POS || v == NEG; }

struct GridInfo {
  int32_t lo[3], hi[3];
  int bc[3];
  int D;
  double dx;

  LSG_HD bool bounded(int d) const { return bc[d] == REFLECTIVE || bc[d] == PERIODIC; }
  // Where a single step from inside [lo, hi] lands when it leaves the
  // axis below lo and above hi: one wrap, mirror or clamp is the whole of
  // map() then (an axis of one point maps everything to it).  With these,
  // map(d, x + s) = stepTo(x + s, ...) for x in [lo, hi] and s = +-1.
  LSG_HD void stepTargets(int d, int32_t &below, int32_t &above) const {
    if (hi[d] == lo[d]) {
      below = above = lo[d];
    } else if (bc[d] == PERIODIC) {
      below = hi[d];
      above = lo[d];
    } else {
      below = (bc[d] == REFLECTIVE || bc[d] == POS_INFINITE) ? lo[d] + 1 : lo[d];
      above = (bc[d] == REFLECTIVE || bc[d] == NEG_INFINITE) ? hi[d] - 1 : hi[d];
    }
  }
  LSG_HD static int32_t stepTo(int64_t y, int32_t lo_, int32_t hi_, int32_t below, int32_t above) {
    return y < lo_ ? below : (y > hi_ ? above : (int32_t)y);
  }
  // periodic wrap / mirror about the boundary point / clamp (same convention
  // as the CPU restatement).  A stencil step leaves the grid by a cell or
  // two: one period or one mirror brings it back, so those come first and
  // the division and the repeated mirrors only run for points further out.
  LSG_HD int32_t map(int d, int64_t c) const {
    if (c >= lo[d] && c <= hi[d]) return (int32_t)c;
    const int64_t l = lo[d], h = hi[d];
    if (bc[d] == PERIODIC) {
      const int64_t L = h - l + 1;
      if (c > h && c - L <= h) return (int32_t)(c - L);
      if (c < l && c + L >= l) return (int32_t)(c + L);
      int64_t m = (c - l) % L;
      if (m < 0) m += L;
      return (int32_t)(l + m);
    }
    const bool reflLo = bc[d] == REFLECTIVE || bc[d] == POS_INFINITE;
    const bool reflHi = bc[d] == REFLECTIVE || bc[d] == NEG_INFINITE;
    c = (c < l) ? (reflLo ? 2 * l - c : l) : (reflHi ? 2 * h - c : h);
    if (c >= l && c <= h) return (int32_t)c;
#if defined(__clang__)
#pragma unroll 1
#endif
    for (int it = 1; it < 8; ++it) {
      if (c < l) c = reflLo ? 2 * l - c : l;
      else if (c > h) c = reflHi ? 2 * h - c : h;
      else break;
    }
    if (c < l) c = l;
    if (c > h) c = h;
    return (int32_t)c;
  }
};

// Unsigned 32-bit division by a run-time constant d >= 1 as a multiply-high
// and a shift (Granlund & Montgomery 1994, round-up variant): exact for
// every 32-bit dividend.
struct FastDiv {
  uint32_t d, m;
  int s;
  LSG_HD void init(uint32_t dd) {
    d = dd;
    s = 0;
    while ((1ull << s) < (uint64_t)dd) ++s;
    m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << s) - dd)) / dd + 1);
  }
  LSG_HD uint32_t div(uint32_t n) const {
    const uint64_t t = ((uint64_t)n * m) >> 32;
    return (uint32_t)((t + n) >> s);
  }
};

// Dense bitmap frame of a band: rows of ext[0] bits padded to whole 64-bit
// words (wpr words per row), rows ordered y fastest then z, no other
// padding.  Frame position p = (z * ext[1] + y) * 64 wpr + x.
struct Frame {
  int32_t o[3];    // absolute index of frame-relative 0 per dim
  int32_t ext[3];  // valid extent per dim (ext[2] == 1 in 2D)
  int wpr;         // words per row
  int base;        // 1 if entry 0 (grid minimum) lies outside the frame
  int64_t nwords;  // words in the bitmap (one zero pad word follows)
  FastDiv divW;    // division by wpr (word -> row)
  FastDiv divY;    // division by ext[1] (row -> z)
  // Derived fields from o/ext/base; returns the number of bits.
  LSG_HD int64_t finish() {
    wpr = (ext[0] + 63) >> 6;
    nwords = (int64_t)wpr * ext[1] * ext[2];
    divW.init((uint32_t)wpr);
    divY.init((uint32_t)ext[1]);
    return nwords * 64;
  }
  LSG_HD uint32_t rowBits() const { return (uint32_t)wpr << 6; }
  LSG_HD uint32_t encode(int32_t xr, int32_t yr, int32_t zr) const {
    return ((uint32_t)zr * (uint32_t)ext[1] + (uint32_t)yr) * rowBits() + (uint32_t)xr;
  }
  LSG_HD void decode(uint32_t p, int32_t &xr, int32_t &yr, int32_t &zr) const {
    const uint32_t row = divW.div(p >> 6);
    xr = (int32_t)(p - row * rowBits());
    const uint32_t z = divY.div(row);
    zr = (int32_t)z;
    yr = (int32_t)(row - z * (uint32_t)ext[1]);
  }
  // Row (y, z) and word-in-row kw of bitmap word w.
  LSG_HD void wordRow(uint32_t w, int32_t &kw, int32_t &yr, int32_t &zr) const {
    const uint32_t row = divW.div(w);
    kw = (int32_t)(w - row * (uint32_t)wpr);
    const uint32_t z = divY.div(row);
    zr = (int32_t)z;
    yr = (int32_t)(row - z * (uint32_t)ext[1]);
  }
};

// Read-only view of a device band.
// Frame position of an entry that is not in the frame: the grid minimum
// point of a band whose frame does not contain it (entry 0 of the band).
constexpr uint32_t POS_NONE = 0xFFFFFFFFu;

struct BandView {
  const uint64_t *bits;  // nwords + 1
  const uint32_t *pre;   // nwords + 1 (exclusive prefix of popcounts, in-frame entries)
  const double *val;     // entries
  const uint32_t *pos;   // entries (frame bit index; POS_NONE for the grid-minimum entry)
  Frame f;
  GridInfo g;
  uint32_t rb;           // entries before the band's first in-frame entry (its grid-minimum entry, if any)

  // Absolute coordinates of entry/candidate position p.
  LSG_HD void coords(uint32_t p, int32_t c[3]) const {
    int32_t xr, yr, zr;
    f.decode(p, xr, yr, zr);
    c[0] = xr + f.o[0];
    c[1] = yr + f.o[1];
    c[2] = zr + f.o[2];
  }

  // A lookup is split in two so that many of them can be issued together:
  // locate() finds the bitmap word/bit of an absolute, in-grid (already
  // mapped) point and whether the point lies inside the frame; valueAt()
  // turns the loaded word and prefix into the value.  Outside the frame the
  // value is that of the run holding the nearest frame position before it.
  struct Loc {
    uint32_t w, bt;
    bool inside;
  };
  LSG_HD Loc locate(int32_t x, int32_t y, int32_t z) const {
    const int64_t xr = (int64_t)x - f.o[0];
    const int64_t yr = (int64_t)y - f.o[1];
    const int64_t zr = (int64_t)z - f.o[2];
    uint64_t lin;
    bool inside = false;
    if (zr < 0) {
      lin = 0;
    } else if (zr >= f.ext[2]) {
      lin = (uint64_t)f.nwords << 6;
    } else {
      uint64_t yc, xc;
      if (yr < 0) { yc = 0; xc = 0; }
      else if (yr >= f.ext[1]) { yc = (uint64_t)f.ext[1]; xc = 0; }
      else {
        yc = (uint64_t)yr;
        if (xr < 0) xc = 0;
        else if (xr >= f.ext[0]) xc = f.rowBits();
        else { xc = (uint64_t)xr; inside = true; }
      }
      lin = ((uint64_t)zr * (uint64_t)f.ext[1] + yc) * f.rowBits() + xc;
    }
    Loc l;
    l.w = (uint32_t)(lin >> 6);
    l.bt = (uint32_t)(lin & 63);
    l.inside = inside;
    return l;
  }
  LSG_HD uint32_t entryOf(const Loc &l, uint64_t word, uint32_t prefix) const {
    const uint32_t rank = rb + prefix + (uint32_t)__builtin_popcountll(word & ((l.bt == 0) ? 0ull : (~0ull >> (64 - l.bt))));
    if (l.inside && ((word >> l.bt) & 1ull)) return rank;
    return rank > 0 ? rank - 1 : 0;
  }
  // Values at N located points: all word/prefix loads, then all value loads.
  template <int N> LSG_HD void values(const Loc *l, double *out) const {
    uint64_t wd[N];
    uint32_t pr[N];
    LSG_UNROLL
    for (int k = 0; k < N; ++k) {
      wd[k] = bits[l[k].w];
      pr[k] = pre[l[k].w];
    }
    LSG_UNROLL
    for (int k = 0; k < N; ++k) out[k] = val[entryOf(l[k], wd[k], pr[k])];
  }
  LSG_HD double value(int32_t x, int32_t y, int32_t z) const {
    const Loc l = locate(x, y, z);
    return val[entryOf(l, bits[l.w], pre[l.w])];
  }
  // Value at centre c displaced by off along dimension d (boundary mapped).
  LSG_HD double at(const int32_t c[3], int d, int off) const {
    int32_t p[3] = {c[0], c[1], c[2]};
    p[d] = g.map(d, (int64_t)c[d] + off);
    return value(p[0], p[1], p[2]);
  }
};

// Parameters of the rate (Hamiltonian) pass.
struct RateParams {
  int spatial;
  int order;
  int calcNormals;
  int checkDissipation;
  double cfl;
  double alphaFactor;
  // built-in velocity
  int velKind, velDir, velK;
  double velB, velAPos, velANeg;
  double vec[3];
  double velAlpha[3];  // getDissipationAlpha of the built-in functor (LLFA)
  // learned weno5 weights: network parameters as bf16 (WENO5 learned only)
  const uint16_t *mlp;
  // host-evaluated velocities: scalar and vector (4 doubles) per record
  // and material, hostRec records per active point in active-list order
  // (the scheme's call, then the LF global-alpha call or the 3^D LLF box);
  // null for the built-in functor
  const double *hostVel;
  int hostRec;
  // stencil local Lax-Friedrichs with host-evaluated velocities: the
  // scalar velocities at the perturbed box normals, 2 D per box point and
  // 2 D 3^D per (active point, material), active point major
  const double *hostSllf;
  // learned weno5: the network of a wave's points on the matrix cores (1)
  // or per lane on the vector ALUs (0)
  int learnedWave;
  // several level sets: built-in functor of material m where matSet[m]
  // (the others use the functor above); kind, dir, k, b, aPos, aNeg, vec
  int matSet[8];
  int adaptive;           // setAdaptiveTimeStepping (several level sets)
  double adaptiveFactor;  // 1 / subdivisions
  int matKind[8], matDir[8], matK[8];
  double matB[8], matAPos[8], matANeg[8], matVec[8][3], matAlpha[8][3];
  // stencil local Lax-Friedrichs: cbrt(DBL_EPSILON), computed on the host
  // (the base of the normal perturbation, lsStencilLocalLaxFriedrichsScalar.hpp:44-45)
  double sllfBaseEps;
  // saveAdvectionVelocities: the scheme's dissipation per point (one level
  // set: per band entry; several: per entry and list slot), else null
  double *dissOut;
  // ignoreVoids: the void marker per point id and the point id of every
  // band entry; a void point gets no scheme call (lsAdvect.hpp:510-528)
  const double *voidMarks;
  const uint32_t *voidRank;
  // batches: per level set the local alpha maxima of the rate pass (LLF,
  // LLLF) and the global alphas found before it (LF), 3 bit patterns of
  // non-negative doubles each
  unsigned long long *segAlpha;
  unsigned long long *segGA;
};

// Which point a band operation's result takes its point data from.
struct PdSrcParams {
  int kind;           // SrcKind
  double totalLimit;  // Expand's width / 2
  int boolKind;       // BoolKind of a boolean
};

// Learned weno5 network: 5 -> 32 -> 32 -> 3 (GELU, softmax); parameter
// layout W1[32][5], b1[32], W2[32][32], b2[32], W3[3][32], b3[3].
constexpr int MLP_IN = 5, MLP_H = 32, MLP_OUT = 3;
constexpr int MLP_W1 = 0, MLP_B1 = 160, MLP_W2 = 192, MLP_B2 = 1216, MLP_W3 = 1248, MLP_B3 = 1344;
constexpr int MLP_NPARAM = 1347;
constexpr double MLP_KINK_RATIO = 10.0;

// fp32 -> bf16 bits, round to nearest even (finite inputs; nan stays nan).
LSG_HD inline uint16_t bf16Bits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0x007fffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
LSG_HD inline float bf16Float(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Pointwise rules of the band operations.
enum OpKind : int {
  OP_EXPAND = 0,    // Expand cycle (limit, totalLimit)
  OP_REBUILD = 1,   // Advect::rebuildLS
  OP_REDUCE = 2,    // Reduce(width)
  OP_PRUNE = 3,     // Prune
  OP_PRUNE_RSZ = 4, // Prune with removeStrayZeros
  OP_UPDATE = 5,    // Reduce(1) + value -= dt * rate
  OP_CHECK = 6,
  OP_UPDATE_MM = 7, // Reduce(1) + update through the rate list of a point (several level sets)
  OP_EXPAND2 = 8,   // two Expand cycles (limits a, then a2; totalLimit b)
  OP_UPDATE_REBUILD = 9  // OP_UPDATE then Advect::rebuildLS (cutoff a) as one star operation
};
enum BoolKind : int { B_INTERSECT = 0, B_UNION = 1, B_RELCOMP = 2, B_COMBINE = 3 };

struct OpParams {
  int kind;
  double a, b;  // expand: limit, totalLimit; reduce: valueLimit; rebuild: cutoff; combine: wTarget, wSource
  const double *rates;
  const double *dt;  // device pointer (update)
  int nmat;          // OP_UPDATE_MM: rate-list slots per entry (level sets)
  double a2;         // OP_EXPAND2: limit of the second cycle
};

}  // namespace lsg
