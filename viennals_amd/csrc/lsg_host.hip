// Host orchestration and C ABI of the MI355X level-set advection library.
//
// A level set (lsg_domain) owns device buffers for its canonical run-start
// list and the frame bitmap index.  All work is enqueued on one HIP stream;
// an advection step runs entirely on the device and the host synchronises
// once per step to read the advanced time (Advect::apply's loop,
// lsAdvect.hpp:1029-1042), the overflow flag and the band bounding box.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/viennals_gpu.h"
#include "lsg_device.hpp"
#include "lsg_geometry.hpp"
#include "lsg_kernels.hpp"
#include "lsg_learned.hpp"

namespace lsg {

static thread_local std::string g_err;
static int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}
#define LSG_HIP(x)                                                                 \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(-100, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

static constexpr int MARGIN = 24;        // frame margin (cells) in infinite dimensions
static constexpr int REFRAME_SLACK = 12; // re-frame when the band comes closer to the frame edge

// Scratch of lsg_expand_noalloc, which allocates nothing: a block of device
// scalars (flags, launch counter), the band statistics and the tile status
// words of the single-pass star kernel for frames of up to
// NOALLOC_MAX_WORDS bitmap words.  Zero at load: tile status words then
// read as "not ready" and the first launch tag is 1.
static constexpr int64_t NOALLOC_MAX_WORDS = (int64_t)1 << 22;  // 2^28 frame points
static constexpr size_t NOALLOC_SC_BYTES = (sizeof(DevScalars) + 255) / 256 * 256;
static constexpr size_t NOALLOC_STAT_BYTES = 256;
static constexpr size_t NOALLOC_TILES = (size_t)(NOALLOC_MAX_WORDS / BAND_TILE_WORDS) + 2;
static constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_SC_BYTES + NOALLOC_STAT_BYTES + NOALLOC_TILES * 8;
static_assert(NOALLOC_SCRATCH_BYTES <= ((size_t)2 << 20), "no-allocation scratch stays near 1 MiB");

}  // namespace lsg

static __device__ __align__(256) unsigned char scratch[lsg::NOALLOC_SCRATCH_BYTES];

namespace lsg {

struct WsStore;
struct Ctx {
  bool init = false;
  int dev = -1;  // the device current at the library's first call; every stream and buffer is on it
  hipStream_t s = nullptr;
  DevScalars *sc = nullptr;
  uint32_t *stat = nullptr;  // [0] count, [1..6] bbox (int32)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  unsigned char *scratch = nullptr;  // device address of `scratch` on dev
};
static Ctx g_ctx;
static WsStore *g_ws = nullptr;

static int ctxInit() {
  if (g_ctx.init) return 0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-101, "no HIP device available");
  LSG_HIP(hipGetDevice(&g_ctx.dev));
  // the array's name is a host-side stand-in: kernels get this address
  LSG_HIP(hipGetSymbolAddress((void **)&g_ctx.scratch, HIP_SYMBOL(scratch)));
  LSG_HIP(hipStreamCreateWithFlags(&g_ctx.s, hipStreamNonBlocking));
  LSG_HIP(hipMalloc(&g_ctx.sc, sizeof(DevScalars)));
  LSG_HIP(hipMemset(g_ctx.sc, 0, sizeof(DevScalars)));
  LSG_HIP(hipMalloc(&g_ctx.stat, 8 * sizeof(uint32_t)));
  LSG_HIP(hipEventCreate(&g_ctx.ev0));
  LSG_HIP(hipEventCreate(&g_ctx.ev1));
  g_ctx.init = true;
  return 0;
}


// Device band storage sized for a frame (words) and an entry capacity.
struct Store {
  BandBufs b{};
  int64_t nwAlloc = 0;
  void release() {
    if (b.bits) (void)hipFree(b.bits);
    if (b.pre) (void)hipFree(b.pre);
    if (b.val) (void)hipFree(b.val);
    if (b.pos) (void)hipFree(b.pos);
    if (b.n) (void)hipFree(b.n);
    b = BandBufs{};
    nwAlloc = 0;
  }
  int alloc(int64_t nw, size_t cap) {
    release();
    LSG_HIP(hipMalloc(&b.bits, (nw + 1) * sizeof(uint64_t)));
    LSG_HIP(hipMalloc(&b.pre, (nw + 1) * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&b.val, cap * sizeof(double)));
    LSG_HIP(hipMalloc(&b.pos, cap * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&b.n, sizeof(uint32_t)));
    b.cap = cap;
    nwAlloc = nw;
    return 0;
  }
};

struct WsStore {
  Workspace w{};
  int64_t nwAlloc = 0;
  void release() {
    (void)hipFree(w.cbits);
    (void)hipFree(w.cbits2);
    (void)hipFree(w.cpre);
    (void)hipFree(w.cpos);
    (void)hipFree(w.out);
    (void)hipFree(w.ks);
    (void)hipFree(w.keepBits);
    (void)hipFree(w.partials);
    (void)hipFree(w.ncand);
    (void)hipFree(w.cseg);
    (void)hipFree(w.tileStat);
    w = Workspace{};
    nwAlloc = 0;
  }
  int alloc(int64_t nw, size_t ccap) {
    release();
    // the candidate list kernel addresses its slots with 32-bit byte offsets
    if (ccap >= ((size_t)1 << 30)) return fail(-12, "band capacity exceeds 2^30 entries");
    LSG_HIP(hipMalloc(&w.cbits, (nw + 1) * sizeof(uint64_t)));
    LSG_HIP(hipMalloc(&w.cbits2, (nw + 1) * sizeof(uint64_t)));
    LSG_HIP(hipMalloc(&w.cpre, (nw + 1) * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&w.cpos, ccap * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&w.out, ccap * sizeof(double)));
    LSG_HIP(hipMalloc(&w.ks, (ccap + 1) * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&w.keepBits, (ccap / 64 + 2) * sizeof(uint64_t)));
    LSG_HIP(hipMalloc(&w.partials, (SCAN_BLOCKS_MAX + 1) * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&w.ncand, sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&w.cseg, ccap * sizeof(uint16_t)));
    // tile status words start as "not ready" (launch tag 0 is never used)
    const size_t ntile = (size_t)(nw / BAND_TILE_WORDS) + 2;
    LSG_HIP(hipMalloc(&w.tileStat, ntile * sizeof(uint64_t)));
    LSG_HIP(hipMemset(w.tileStat, 0, ntile * sizeof(uint64_t)));
    w.ccap = ccap;
    nwAlloc = nw;
    return 0;
  }
};

}  // namespace lsg

using namespace lsg;

// One array of Domain::getPointData (lsDomain.hpp:44): a scalar (1
// component) or vector (3 components) value per defined point, in point-id
// order, resident on the device; `alt` receives the next translation.
struct PdArray {
  std::string label;
  int comps = 1;
  double *buf = nullptr, *alt = nullptr;
  size_t cap = 0;  // tuples
};

struct lsg_domain {
  int D = 3;
  GridInfo g{};
  int32_t umin[3] = {0, 0, 0}, umax[3] = {0, 0, 0};  // the grid bounds as given (kept for serialisation)
  Frame f{};
  int width = 1;
  Store cur, tmp;  // current band and the output buffer of the next op
  uint32_t hostN = 0;
  bool framed = false;
  std::vector<PdArray> pd;    // point data, insertion order
  uint32_t *pdLen = nullptr;  // device: tuples in every array

  void pdRelease() {
    for (PdArray &a : pd) {
      if (a.buf) (void)hipFree(a.buf);
      if (a.alt) (void)hipFree(a.alt);
    }
    pd.clear();
  }
  ~lsg_domain() {
    cur.release();
    tmp.release();
    pdRelease();
    if (pdLen) (void)hipFree(pdLen);
  }
};

namespace lsg {

// ------------------------------------------------------------- frames ---

static bool gridMinInFrame(const lsg_domain *d, const Frame &f) {
  for (int k = 0; k < d->D; ++k)
    if (f.o[k] != d->g.lo[k]) return false;
  return true;
}

// Frame covering the given absolute bounding box (bounded dimensions span
// the grid; infinite dimensions get MARGIN cells on each side).
static int makeFrame(const lsg_domain *d, const int32_t *bbmin, const int32_t *bbmax, Frame &f) {
  f = Frame{};
  for (int k = 0; k < 3; ++k) {
    if (k >= d->D) {
      f.o[k] = 0;
      f.ext[k] = 1;
      continue;
    }
    if (d->g.bounded(k)) {
      f.o[k] = d->g.lo[k];
      f.ext[k] = d->g.hi[k] - d->g.lo[k] + 1;
    } else {
      int64_t lo = (int64_t)bbmin[k] - MARGIN, hi = (int64_t)bbmax[k] + MARGIN;
      lo = std::max<int64_t>(lo, d->g.lo[k]);
      hi = std::min<int64_t>(hi, d->g.hi[k]);
      f.o[k] = (int32_t)lo;
      f.ext[k] = (int32_t)(hi - lo + 1);
    }
  }
  const int64_t bitsTotal = (int64_t)((f.ext[0] + 63) >> 6) * 64 * f.ext[1] * f.ext[2];
  if (bitsTotal >= ((int64_t)1 << 32) - 64)
    return fail(-10, "level set frame exceeds 2^32 grid points; reduce the domain or band extent");
  f.finish();
  f.base = gridMinInFrame(d, f) ? 0 : 1;
  return 0;
}

// Initial entry capacity of a band.  LSG_TEST_CAPACITY (tests only) sets a
// small capacity to exercise the overflow-retry path.
static size_t capFor(size_t n) {
  if (const char *t = getenv("LSG_TEST_CAPACITY")) {
    size_t c = (size_t)strtoull(t, nullptr, 10);
    if (c > 0) return std::max<size_t>(c, n + 64);
  }
  return std::max<size_t>((size_t)1 << 20, n * 8 + 4096);
}

static int ensureStores(lsg_domain *d, size_t cap) {
  const int64_t nw = d->f.nwords;
  if (d->cur.b.cap < cap || d->cur.nwAlloc < nw) {
    // preserve current contents
    Store n;
    if (int e = n.alloc(std::max(nw, d->cur.nwAlloc), std::max(cap, d->cur.b.cap))) return e;
    if (d->cur.b.val) {
      LSG_HIP(hipMemcpyAsync(n.b.val, d->cur.b.val, d->cur.b.cap * sizeof(double), hipMemcpyDeviceToDevice, g_ctx.s));
      LSG_HIP(hipMemcpyAsync(n.b.pos, d->cur.b.pos, d->cur.b.cap * sizeof(uint32_t), hipMemcpyDeviceToDevice, g_ctx.s));
      LSG_HIP(hipMemcpyAsync(n.b.bits, d->cur.b.bits, (d->cur.nwAlloc + 1) * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                             g_ctx.s));
      LSG_HIP(hipMemcpyAsync(n.b.pre, d->cur.b.pre, (d->cur.nwAlloc + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                             g_ctx.s));
      LSG_HIP(hipMemcpyAsync(n.b.n, d->cur.b.n, sizeof(uint32_t), hipMemcpyDeviceToDevice, g_ctx.s));
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
    }
    d->cur.release();
    d->cur = n;
    n.b = BandBufs{};
  }
  if (d->tmp.b.cap < d->cur.b.cap || d->tmp.nwAlloc < d->cur.nwAlloc)
    if (int e = d->tmp.alloc(d->cur.nwAlloc, d->cur.b.cap)) return e;
  if (!g_ws) g_ws = new WsStore();
  if (g_ws->w.ccap < d->cur.b.cap || g_ws->nwAlloc < d->cur.nwAlloc) {
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    if (int e = g_ws->alloc(std::max(g_ws->nwAlloc, d->cur.nwAlloc), std::max(g_ws->w.ccap, d->cur.b.cap))) return e;
  }
  return 0;
}

// Allocate only the current-band storage (copies kept for retries).
static int ensureCur(lsg_domain *d, int64_t nw, size_t cap) {
  if (d->cur.b.cap >= cap && d->cur.nwAlloc >= nw) return 0;
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return d->cur.alloc(std::max(nw, d->cur.nwAlloc), std::max(cap, d->cur.b.cap));
}

// Upload a canonical run-start list (entry 0 = grid minimum point).
static int uploadRuns(lsg_domain *d, const std::vector<int32_t> &idx, const std::vector<double> &val, int width,
                      const Frame *forced = nullptr) {
  if (int e = ctxInit()) return e;
  const int D = d->D;
  const size_t n = val.size();
  if (n == 0) return fail(-11, "empty run list");
  for (int k = 0; k < D; ++k)
    if (idx[k] != d->g.lo[k]) return fail(-12, "run list must start at the grid minimum point");
  int32_t bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
  bool any = false;
  for (size_t i = 1; i < n; ++i) {
    for (int k = 0; k < D; ++k) {
      int32_t c = idx[i * D + k];
      if (c < d->g.lo[k] || c > d->g.hi[k]) return fail(-13, "run list point outside the grid");
      if (!any || c < bbmin[k]) bbmin[k] = c;
      if (!any || c > bbmax[k]) bbmax[k] = c;
    }
    any = true;
  }
  if (!any)
    for (int k = 0; k < D; ++k) bbmin[k] = bbmax[k] = (d->g.bounded(k) ? d->g.lo[k] : 0);
  Frame f;
  if (forced) f = *forced;
  else if (int e = makeFrame(d, bbmin, bbmax, f)) return e;
  d->f = f;
  d->framed = true;
  // host-side encode
  std::vector<uint32_t> pos(n, 0xFFFFFFFFu);
  std::vector<uint64_t> bits(f.nwords + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    if (i == 0 && f.base) continue;
    int32_t r[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) r[k] = idx[i * D + k] - f.o[k];
    if (r[0] < 0 || r[0] >= f.ext[0] || r[1] < 0 || r[1] >= f.ext[1] || r[2] < 0 || r[2] >= f.ext[2])
      return fail(-14, "internal: entry outside frame");
    uint32_t p = f.encode(r[0], r[1], r[2]);
    if (i > (size_t)f.base && p <= pos[i - 1]) return fail(-15, "run list is not strictly sorted in HRLE order");
    pos[i] = p;
    bits[p >> 6] |= 1ull << (p & 63);
  }
  std::vector<uint32_t> pre(f.nwords + 1, 0);
  uint32_t acc = 0;
  for (int64_t w = 0; w <= f.nwords; ++w) {
    pre[w] = acc;
    if (w < f.nwords) acc += (uint32_t)__builtin_popcountll(bits[w]);
  }
  if (int e = ensureStores(d, capFor(n))) return e;
  BandBufs &b = d->cur.b;
  uint32_t nn = (uint32_t)n;
  LSG_HIP(hipMemcpyAsync(b.val, val.data(), n * sizeof(double), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(b.pos, pos.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(b.bits, bits.data(), (f.nwords + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(b.pre, pre.data(), (f.nwords + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(b.n, &nn, sizeof(uint32_t), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  d->hostN = nn;
  d->width = width;
  return 0;
}

static int downloadRuns(lsg_domain *d, std::vector<int32_t> &idx, std::vector<double> &val) {
  if (int e = ctxInit()) return e;
  uint32_t n = 0;
  LSG_HIP(hipMemcpyAsync(&n, d->cur.b.n, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  std::vector<uint32_t> pos(n);
  val.resize(n);
  LSG_HIP(hipMemcpyAsync(val.data(), d->cur.b.val, n * sizeof(double), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(pos.data(), d->cur.b.pos, n * sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  const int D = d->D;
  idx.resize((size_t)n * D);
  for (uint32_t i = 0; i < n; ++i) {
    if (i == 0 && d->f.base) {
      for (int k = 0; k < D; ++k) idx[k] = d->g.lo[k];
      continue;
    }
    int32_t r[3];
    d->f.decode(pos[i], r[0], r[1], r[2]);
    for (int k = 0; k < D; ++k) idx[(size_t)i * D + k] = r[k] + d->f.o[k];
  }
  d->hostN = n;
  return 0;
}

static bool idxLessD(int D, const int32_t *a, const int32_t *b) {
  for (int k = D - 1; k >= 0; --k)
    if (a[k] != b[k]) return a[k] < b[k];
  return false;
}

// Canonical run list from defined points: the run after a defined point is
// negative iff the point's value is negative; the run before the first
// point takes its sign (the role of HRLE's FillDomainWithSignedDistance
// behind Domain::insertPoints, lsDomain.hpp:135-138).
static void runsFromPoints(const lsg_domain *d, std::vector<int32_t> pidx, std::vector<double> pval,
                           std::vector<int32_t> &idx, std::vector<double> &val) {
  const int D = d->D;
  {  // points outside the grid are dropped (as in the CPU restatement)
    size_t w = 0;
    for (size_t i = 0; i < pval.size(); ++i) {
      bool in = true;
      for (int k = 0; k < D; ++k) in = in && pidx[i * D + k] >= d->g.lo[k] && pidx[i * D + k] <= d->g.hi[k];
      if (!in) continue;
      for (int k = 0; k < D; ++k) pidx[w * D + k] = pidx[i * D + k];
      pval[w++] = pval[i];
    }
    pidx.resize(w * D);
    pval.resize(w);
  }
  const size_t n = pval.size();
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return idxLessD(D, &pidx[a * D], &pidx[b * D]); });
  idx.clear();
  val.clear();
  int32_t gmin[3] = {d->g.lo[0], d->g.lo[1], d->g.lo[2]};
  for (int k = 0; k < D; ++k) idx.push_back(gmin[k]);
  val.push_back(n == 0 ? POS : (pval[order[0]] < 0 ? NEG : POS));
  std::vector<int32_t> cidx;
  std::vector<double> cval;
  for (size_t j = 0; j < n; ++j) {
    const int32_t *p = &pidx[order[j] * D];
    if (j > 0 && !idxLessD(D, &pidx[order[j - 1] * D], p)) continue;  // duplicate
    bool isMin = true;
    for (int k = 0; k < D; ++k) isMin = isMin && p[k] == gmin[k];
    if (isMin) val[0] = pval[order[j]];
    else {
      for (int k = 0; k < D; ++k) idx.push_back(p[k]);
      val.push_back(pval[order[j]]);
    }
    int32_t s[3] = {p[0], D > 1 ? p[1] : 0, D > 2 ? p[2] : 0};
    int k = 0;
    for (; k < D; ++k) {
      if (s[k] < d->g.hi[k]) {
        s[k]++;
        break;
      }
      s[k] = d->g.lo[k];
    }
    if (k == D) continue;
    bool beforeNext = (j + 1 >= n) || idxLessD(D, s, &pidx[order[j + 1] * D]);
    if (beforeNext) {
      for (int q = 0; q < D; ++q) idx.push_back(s[q]);
      val.push_back(pval[order[j]] < 0 ? NEG : POS);
    }
  }
  // compress to the canonical form
  std::vector<int32_t> oi;
  std::vector<double> ov;
  for (size_t i = 0; i < val.size(); ++i) {
    if (i == 0 || !isUndef(val[i]) || val[i] != val[i - 1]) {
      for (int q = 0; q < D; ++q) oi.push_back(idx[i * D + q]);
      ov.push_back(val[i]);
    }
  }
  idx.swap(oi);
  val.swap(ov);
}

// ------------------------------------------------------------- ops ----

static void swapCur(lsg_domain *d) { std::swap(d->cur, d->tmp); }

// Event pairs around the evaluation kernel of star operations, taken from a
// pool while an Advect with profiling on enqueues its step.
struct StarTiming {
  hipEvent_t *ev = nullptr;
  int cap = 0, used = 0;
  bool external = false;
};
static StarTiming *g_starTiming = nullptr;   // one-cycle star operations
static StarTiming *g_star2Timing = nullptr;  // two Expand cycles at once

static void takeEvents(StarTiming *t, hipEvent_t &e0, hipEvent_t &e1, bool &external) {
  if (t && t->used + 2 <= t->cap) {
    e0 = t->ev[t->used];
    e1 = t->ev[t->used + 1];
    t->used += 2;
    external = t->external;
  }
}
static void opStar(lsg_domain *d, const OpParams &op) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool external = false;
  takeEvents(g_starTiming, e0, e1, external);
  launchStarOp(g_ctx.s, d->D, d->cur.b, d->tmp.b, d->f, d->g, g_ws->w, op, g_ctx.sc, e0, e1, external);
  swapCur(d);
}
static void opStar2(lsg_domain *d, const OpParams &op) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool external = false;
  takeEvents(g_star2Timing, e0, e1, external);
  launchStarOp2(g_ctx.s, d->D, d->cur.b, d->tmp.b, d->f, d->g, g_ws->w, op, g_ctx.sc, e0, e1, external);
  swapCur(d);
}
static void opSelf(lsg_domain *d, const OpParams &op) {
  launchSelfOp(g_ctx.s, d->D, d->cur.b, d->tmp.b, d->f, d->g, g_ws->w, op, g_ctx.sc);
  swapCur(d);
}

// ------------------------------------------------------------ point data ---
// PointData::translateFromMultiData after every band operation (all of them
// update point data by default: lsExpand.hpp:27, lsReduce.hpp:19,
// lsPrune.hpp:19, lsBooleanOperation.hpp:56; Advect::rebuildLS only with
// updatePointData, lsAdvect.hpp:75,262).  The input band of the operation is
// the domain's `tmp` store right after it ran.
struct PdWork {
  uint32_t *rankIn = nullptr, *rankB = nullptr, *dlist = nullptr, *src = nullptr, *entry = nullptr;
  uint8_t *side = nullptr;
  uint32_t *ndef = nullptr, *nPush = nullptr;
  size_t cap = 0;
  int ensure(size_t c) {
    if (c <= cap) return 0;
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    for (void *p : {(void *)rankIn, (void *)rankB, (void *)dlist, (void *)src, (void *)entry, (void *)side})
      if (p) (void)hipFree(p);
    LSG_HIP(hipMalloc(&rankIn, c * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&rankB, c * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&dlist, c * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&src, c * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&entry, c * sizeof(uint32_t)));
    LSG_HIP(hipMalloc(&side, c));
    if (!ndef) LSG_HIP(hipMalloc(&ndef, sizeof(uint32_t)));
    if (!nPush) LSG_HIP(hipMalloc(&nPush, sizeof(uint32_t)));
    cap = c;
    return 0;
  }
};
static PdWork *g_pd = nullptr;

static int pdEnsureArray(PdArray &a, size_t cap, const uint32_t *len) {
  if (a.cap >= cap) return 0;
  double *nb = nullptr, *na = nullptr;
  LSG_HIP(hipMalloc(&nb, cap * a.comps * sizeof(double)));
  LSG_HIP(hipMalloc(&na, cap * a.comps * sizeof(double)));
  if (a.buf) {
    (void)len;
    LSG_HIP(hipMemcpyAsync(nb, a.buf, a.cap * a.comps * sizeof(double), hipMemcpyDeviceToDevice, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    (void)hipFree(a.buf);
  }
  if (a.alt) (void)hipFree(a.alt);
  a.buf = nb;
  a.alt = na;
  a.cap = cap;
  return 0;
}
static int pdEnsure(lsg_domain *d, size_t cap) {
  if (!d->pdLen) {
    LSG_HIP(hipMalloc(&d->pdLen, sizeof(uint32_t)));
    LSG_HIP(hipMemsetAsync(d->pdLen, 0, sizeof(uint32_t), g_ctx.s));
  }
  for (PdArray &a : d->pd)
    if (int e = pdEnsureArray(a, cap, d->pdLen)) return e;
  return 0;
}
static PdArray *pdFind(lsg_domain *d, const char *label, int comps) {
  for (PdArray &a : d->pd)
    if (a.comps == comps && a.label == label) return &a;
  return nullptr;
}
static const PdArray *pdFind(const lsg_domain *d, const std::string &label, int comps) {
  for (const PdArray &a : d->pd)
    if (a.comps == comps && a.label == label) return &a;
  return nullptr;
}

// Sources of the result of the operation that just ran on d (into g_pd).
static int pdSources(lsg_domain *d, const PdSrcParams &p, const lsg_domain *B = nullptr) {
  if (!g_pd) g_pd = new PdWork();
  const size_t cap = std::max(std::max(d->cur.b.cap, d->tmp.b.cap), B ? B->cur.b.cap : (size_t)0);
  if (int e = g_pd->ensure(cap)) return e;
  launchPointIds(g_ctx.s, d->tmp.b, g_ws->w, g_pd->dlist, g_pd->ndef, g_pd->rankIn);
  if (B) launchPointIds(g_ctx.s, B->cur.b, g_ws->w, g_pd->dlist, g_pd->ndef, g_pd->rankB);
  launchPdSources(g_ctx.s, d->D, d->cur.b, d->tmp.b, B ? &B->cur.b : nullptr, d->f, d->g, p, g_ws->w, g_pd->rankIn,
                  B ? g_pd->rankB : nullptr, g_pd->src, g_pd->side, g_pd->entry, g_pd->nPush);
  return 0;
}
// Gather every array through the sources (a boolean's B-sourced points take
// B's array of the same label, else 0: lsBooleanOperation.hpp:146-196).
static int pdApply(lsg_domain *d, const lsg_domain *B = nullptr) {
  if (int e = pdEnsure(d, d->cur.b.cap)) return e;
  for (PdArray &a : d->pd) {
    const PdArray *b = B ? pdFind(B, a.label, a.comps) : nullptr;
    launchPdGather(g_ctx.s, a.comps, a.buf, d->pdLen, b ? b->buf : nullptr, b ? B->pdLen : nullptr, g_pd->src,
                   g_pd->side, g_pd->nPush, d->cur.b.cap, a.alt);
    std::swap(a.buf, a.alt);
  }
  launchCopyU32(g_ctx.s, g_pd->nPush, d->pdLen);
  return 0;
}
static void pdTranslate(lsg_domain *d, const PdSrcParams &p, const lsg_domain *B = nullptr) {
  if (d->pd.empty()) return;
  if (pdSources(d, p, B) == 0) (void)pdApply(d, B);
}
// MarkVoidPoints on the expanded band before the rate pass (lsAdvect.hpp:443-
// 453): "VoidPointMarkers" becomes point data of the level set (one value
// per point of this band) and the rate kernels skip the void points.
struct VoidWork {
  uint32_t *dlist = nullptr, *ndef = nullptr, *rank = nullptr, *par = nullptr, *topQ = nullptr;
  uint8_t *sign = nullptr;
  size_t cap = 0;
  uint64_t nbox = 0;
  int ensure(size_t c, uint64_t N) {
    if (c > cap) {
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      for (void *p : {(void *)dlist, (void *)rank}) if (p) (void)hipFree(p);
      LSG_HIP(hipMalloc(&dlist, c * sizeof(uint32_t)));
      LSG_HIP(hipMalloc(&rank, c * sizeof(uint32_t)));
      cap = c;
    }
    if (N > nbox) {
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      for (void *p : {(void *)par, (void *)sign}) if (p) (void)hipFree(p);
      LSG_HIP(hipMalloc(&par, N * sizeof(uint32_t)));
      LSG_HIP(hipMalloc(&sign, N));
      nbox = N;
    }
    if (!ndef) LSG_HIP(hipMalloc(&ndef, sizeof(uint32_t)));
    if (!topQ) LSG_HIP(hipMalloc(&topQ, sizeof(uint32_t)));
    return 0;
  }
};
static VoidWork *g_void = nullptr;
static void pdInsertEmpty(lsg_domain *d, const char *label) {
  for (PdArray &a : d->pd)
    if (a.comps == 1 && a.label == label) return;
  d->pd.emplace_back();
  d->pd.back().label = label;
}
static int markVoids(lsg_domain *d, RateParams &rp) {
  if (!g_void) g_void = new VoidWork();
  const uint64_t N = (uint64_t)d->f.ext[0] * d->f.ext[1] * d->f.ext[2];
  if (N >= 0xFFFFFFFEull) return fail(-5, "Advect: ignoreVoids needs a frame of fewer than 2^32 points");
  if (int e = g_void->ensure(d->cur.b.cap, N)) return e;
  launchPointIds(g_ctx.s, d->cur.b, g_ws->w, g_void->dlist, g_void->ndef, g_void->rank);
  pdInsertEmpty(d, "VoidPointMarkers");  // MarkVoidPoints::voidPointLabel (lsMarkVoidPoints.hpp:86)
  if (int e = pdEnsure(d, d->cur.b.cap)) return e;
  double *marks = nullptr;
  for (PdArray &a : d->pd)
    if (a.comps == 1 && a.label == "VoidPointMarkers") marks = a.buf;
  launchVoidMarks(g_ctx.s, d->D, d->cur.b, d->f, d->g, g_void->sign, g_void->par, g_void->topQ, g_void->dlist,
                  g_void->ndef, marks);
  launchCopyU32(g_ctx.s, g_void->ndef, d->pdLen);
  rp.voidMarks = marks;
  rp.voidRank = g_void->rank;
  return 0;
}

static void pdClear(lsg_domain *d) {
  if (d->pd.empty()) return;
  (void)hipStreamSynchronize(g_ctx.s);
  d->pdRelease();
}
// dst's point data becomes a copy of src's (Domain::deepCopy, lsDomain.hpp:143-150).
static int pdCopy(lsg_domain *dst, const lsg_domain *src) {
  if (dst->pd.empty() && src->pd.empty()) return 0;
  if (dst->pd.size() != src->pd.size()) {
    (void)hipStreamSynchronize(g_ctx.s);
    dst->pdRelease();
    dst->pd.resize(src->pd.size());
  }
  for (size_t i = 0; i < src->pd.size(); ++i) {
    dst->pd[i].label = src->pd[i].label;
    if (dst->pd[i].comps != src->pd[i].comps) {
      if (dst->pd[i].buf) (void)hipFree(dst->pd[i].buf);
      if (dst->pd[i].alt) (void)hipFree(dst->pd[i].alt);
      dst->pd[i].buf = dst->pd[i].alt = nullptr;
      dst->pd[i].cap = 0;
      dst->pd[i].comps = src->pd[i].comps;
    }
  }
  size_t cap = 0;
  for (const PdArray &a : src->pd) cap = std::max(cap, a.cap);
  if (int e = pdEnsure(dst, cap)) return e;
  for (size_t i = 0; i < src->pd.size(); ++i)
    LSG_HIP(hipMemcpyAsync(dst->pd[i].buf, src->pd[i].buf, src->pd[i].cap * src->pd[i].comps * sizeof(double),
                           hipMemcpyDeviceToDevice, g_ctx.s));
  if (src->pdLen) launchCopyU32(g_ctx.s, src->pdLen, dst->pdLen);
  return 0;
}

// Expand<T,D>::apply (lsExpand.hpp:44-172); the early exit for a band
// without defined points is not taken (its result is then unchanged).
// LSG_EXPAND_PAIRS=1: with the multi-kernel band operations and no point
// data to carry, the cycles run two at a time (OP_EXPAND2), the pairs
// first: the candidates of a pair reach two cells beyond the band, so the
// widest candidate set is the one of the last single cycle or pair, as with
// single cycles.  Off by default: the pair's evaluation reads 31 points per
// candidate and costs what the two operations it replaces cost.
static bool expandPairs() {
  static const bool on = [] {
    const char *e = getenv("LSG_EXPAND_PAIRS");
    return e && e[0] == '1';
  }();
  return on;
}
// The update of a stage and the rebuild after it as one band operation;
// LSG_FUSE_UPDATE=0 keeps them apart.
static bool fuseUpdateRebuild() {
  static const bool on = [] {
    const char *e = getenv("LSG_FUSE_UPDATE");
    return !(e && e[0] == '0');
  }();
  return on;
}
static void doExpand(lsg_domain *d, int width) {
  if (width <= d->width) return;
  const double totalLimit = width * 0.5;
  const int startWidth = d->width;
  const int cycles = width - startWidth;
  const bool pairs = expandPairs() && !singlePassBandOps() && d->pd.empty();
  for (int cyc = 0; cyc < cycles;) {
    OpParams op{};
    op.a = (startWidth + cyc + 1) * 0.5;
    op.b = totalLimit;
    if (pairs && cycles - cyc >= 2) {
      op.kind = OP_EXPAND2;
      op.a2 = (startWidth + cyc + 2) * 0.5;
      opStar2(d, op);
      cyc += 2;
    } else {
      op.kind = OP_EXPAND;
      opStar(d, op);
      pdTranslate(d, PdSrcParams{SRC_EXPAND, totalLimit, 0});
      cyc += 1;
    }
  }
  d->width = width;
}
static void doReduce(lsg_domain *d, int width) {  // lsReduce.hpp:55-134
  if (width >= d->width) return;
  OpParams op{};
  op.kind = OP_REDUCE;
  op.a = width * 0.5;
  opSelf(d, op);
  pdTranslate(d, PdSrcParams{SRC_CENTRE, 0, 0});
  d->width = width;
}
static void doPrune(lsg_domain *d, bool rsz) {  // lsPrune.hpp:74-194
  OpParams op{};
  op.kind = rsz ? OP_PRUNE_RSZ : OP_PRUNE;
  opSelf(d, op);
  pdTranslate(d, PdSrcParams{rsz ? SRC_PRUNE_RSZ : SRC_PRUNE, 0, 0});
  d->width = 2;
}
// Advect::rebuildLS (lsAdvect.hpp:238-429): cutoff 1.0 and final width 2,
// 1.5 and 3 for the stencil local Lax-Friedrichs scheme (:248-256); the
// point data follows when updateData, else the result has none (:262-267,
// 418-427).
static void doRebuild(lsg_domain *d, bool sllf = false, bool updateData = true) {
  OpParams op{};
  op.kind = OP_REBUILD;
  op.a = sllf ? 1.5 : 1.0;
  opStar(d, op);
  if (updateData) pdTranslate(d, PdSrcParams{SRC_REBUILD, 0, 0});
  else pdClear(d);
  d->width = sllf ? 3 : 2;
}
// booleanOpInternal + pruning (lsBooleanOperation.hpp:59-213)
static void doBoolean(lsg_domain *a, const lsg_domain *b, int kind, double wT, double wS,
                      uint32_t *movedDown = nullptr) {
  launchBoolean(g_ctx.s, a->D, a->cur.b, b->cur.b, a->tmp.b, a->f, a->g, g_ws->w, kind, wT, wS, g_ctx.sc, movedDown);
  swapCur(a);
  pdTranslate(a, PdSrcParams{SRC_BOOLEAN, 0, kind}, b);
  doPrune(a, true);
  doPrune(a, false);
}

// A look-back wait of a single-pass band operation ran out (not expected:
// it needs workgroups dispatched out of index order).  The single-pass
// kernels are turned off for the rest of the process and the flag cleared;
// the caller repeats its work from a saved copy with the multi-kernel
// pipeline.
static bool lookbackFailed(const DevScalars &h) {
  if (!h.lookbackFail) return false;
  disableSinglePassBandOps();
  fprintf(stderr, "[viennals_gpu] a single-pass band operation did not complete; repeating with the multi-kernel "
                  "band operations\n");
  (void)hipMemsetAsync(&g_ctx.sc->lookbackFail, 0, sizeof(uint32_t), g_ctx.s);
  (void)hipStreamSynchronize(g_ctx.s);
  return true;
}

static int readOverflow(bool reset) {
  uint32_t ov = 0;
  LSG_HIP(hipMemcpyAsync(&ov, &g_ctx.sc->overflow, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  if (ov && reset) LSG_HIP(hipMemsetAsync(&g_ctx.sc->overflow, 0, sizeof(uint32_t), g_ctx.s));
  return ov ? 1 : 0;
}

// Band statistics: defined count and absolute bounding box of in-frame entries.
static int bandStats(lsg_domain *d, uint32_t &ndef, int32_t bbmin[3], int32_t bbmax[3], bool &any) {
  launchBandStats(g_ctx.s, d->cur.b, d->f, g_ctx.stat, (int32_t *)(g_ctx.stat + 1));
  uint32_t h[7];
  LSG_HIP(hipMemcpyAsync(h, g_ctx.stat, 7 * sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  ndef = h[0];
  const int32_t *bb = (const int32_t *)(h + 1);
  any = bb[0] <= bb[3];
  for (int k = 0; k < 3; ++k) {
    bbmin[k] = bb[k] + d->f.o[k];
    bbmax[k] = bb[3 + k] + d->f.o[k];
  }
  return 0;
}

// Move the frame when the band approached its edge in an infinite dimension.
static int maybeReframe(lsg_domain *d, bool force) {
  uint32_t ndef;
  int32_t bbmin[3], bbmax[3];
  bool any;
  if (int e = bandStats(d, ndef, bbmin, bbmax, any)) return e;
  if (!any) return 0;
  bool need = force;
  for (int k = 0; k < d->D; ++k) {
    if (d->g.bounded(k)) continue;
    if (bbmin[k] - d->f.o[k] < REFRAME_SLACK || (d->f.o[k] + d->f.ext[k] - 1) - bbmax[k] < REFRAME_SLACK)
      need = true;
  }
  if (!need) return 0;
  std::vector<int32_t> idx;
  std::vector<double> val;
  if (int e = downloadRuns(d, idx, val)) return e;
  int w = d->width;
  return uploadRuns(d, idx, val, w);
}

// Bring level sets onto one frame covering all their bands (BooleanOperation
// inputs; the level sets of a multi-material Advect).
static int commonFrameAll(const std::vector<lsg_domain *> &ds) {
  lsg_domain *a = ds[0];
  bool same = true;
  for (lsg_domain *d : ds) {
    if (a->D != d->D) return fail(-20, "level sets of different dimension");
    for (int k = 0; k < 3; ++k)
      if (a->g.lo[k] != d->g.lo[k] || a->g.hi[k] != d->g.hi[k] || a->g.bc[k] != d->g.bc[k])
        return fail(-21, "level sets on different grids");
    if (a->g.dx != d->g.dx) return fail(-21, "level sets on different grids");
    if (!d->framed) return fail(-7, "level set has no data");
    if (std::memcmp(&a->f, &d->f, sizeof(Frame)) != 0) same = false;
  }
  if (same) return 0;
  int32_t mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
  bool any = false;
  for (lsg_domain *d : ds) {
    uint32_t nd;
    int32_t dmin[3], dmax[3];
    bool da;
    if (int e = bandStats(d, nd, dmin, dmax, da)) return e;
    if (!da) continue;
    for (int k = 0; k < 3; ++k) {
      mn[k] = any ? std::min(mn[k], dmin[k]) : dmin[k];
      mx[k] = any ? std::max(mx[k], dmax[k]) : dmax[k];
    }
    any = true;
  }
  Frame f;
  if (int e = makeFrame(a, mn, mx, f)) return e;
  for (lsg_domain *d : ds) {
    std::vector<int32_t> idx;
    std::vector<double> val;
    if (int e = downloadRuns(d, idx, val)) return e;
    if (int e = uploadRuns(d, idx, val, d->width, &f)) return e;
  }
  return 0;
}
static int commonFrame(lsg_domain *a, lsg_domain *b) { return commonFrameAll({a, b}); }

}  // namespace lsg

// =============================================================== Advect ===

struct lsg_advect {
  std::vector<lsg_domain *> levelSets;
  int spatial = 0, temporal = 0;
  double cfl = 0.4999, advectionTime = 0.0, dissipationAlpha = 1.0;
  int singleStep = 0, calcNormals = 1, checkDissipation = 1, ignoreVoids = 0;
  bool velSet = false;
  RateParams rp{};
  double advectedTime = 0, currentTimeStep = -1, pointUpdates = 0;
  unsigned numSteps = 0;
  lsg_domain *init = nullptr;  // RK initial level set
  lsg_domain *save = nullptr;  // band at the start of a step (capacity-overflow retry)
  std::vector<lsg_domain *> lowerSave;  // several level sets: the lower ones at the start of a step
  double *rates = nullptr;
  size_t ratesCap = 0;
  // point data (lsAdvect.hpp:74-75): translated by rebuildLS, and the saved
  // velocities; diss holds the scheme's dissipation per rate slot
  bool updatePointData = true, saveVelocities = false;
  double *diss = nullptr;
  bool profiling = false;
  double rateMs = 0, rateBytes = 0;
  int64_t rateLaunches = 0;
  // rate-kernel timing events (pairs), reused by every step
  std::vector<hipEvent_t> evPool;
  int evUsed = 0;
  bool capturing = false;
  // star-operation evaluation kernel (profiling)
  std::vector<hipEvent_t> starPool, star2Pool;
  lsg::StarTiming starT, star2T;
  double starMs = 0, starBytes = 0, star2Ms = 0, star2Bytes = 0;
  int64_t starLaunches = 0, star2Launches = 0;
  // one captured graph per buffer parity of the step
  struct StepGraph {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    lsg::Store dCur, dTmp, iCur, iTmp;
    int dWidth = 0, iWidth = 0, evUsed = 0, starUsed = 0;
  };
  StepGraph graphs[2];
  int nextGraph = 0;
  bool useGraphs = true;
  double *hostMaxDt = nullptr;  // pinned, read by the step's first kernel
  double *devMaxDt = nullptr;
  lsg::DevScalars *hostSc = nullptr;  // pinned
  uint32_t *hostStat = nullptr;       // pinned
  uint16_t *mlpDev = nullptr;         // learned weno5 parameters (bf16)
  // host-evaluated velocity (VelocityField subclass on the host)
  lsg_scalar_velocity_fn hostScalar = nullptr;
  lsg_vector_velocity_fn hostVector = nullptr;
  void *hostUser = nullptr;
  lsg_dissipation_alpha_fn hostAlpha = nullptr;  // analytical LLF: getDissipationAlpha on the host
  void *hostAlphaUser = nullptr;
  double *velIn = nullptr, *velOut = nullptr;  // device: 7 / 4 (x materials) doubles per active point
  int velMat = 1, velRec = 1;
  uint32_t *dlist = nullptr, *ndef = nullptr;
  size_t velCap = 0;
  std::vector<double> velInH, velOutH;
  // stencil local Lax-Friedrichs: second-round records / results (device)
  double *sllfIn = nullptr, *sllfOut = nullptr;
  size_t sllfCap = 0;
  std::vector<double> sllfInH, sllfOutH;
  int hostErr = 0;
  // Advect::setVelocityUpdateCallback (lsAdvect.hpp:82,966-969)
  lsg_velocity_update_fn velUpdate = nullptr;
  void *velUpdateUser = nullptr;
  void dropGraphs() {
    for (auto &g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = StepGraph{};
    }
  }
  ~lsg_advect() {
    dropGraphs();
    if (init) delete init;
    if (save) delete save;
    for (lsg_domain *d : lowerSave) delete d;
    if (rates) (void)hipFree(rates);
    if (diss) (void)hipFree(diss);
    for (auto e : evPool) (void)hipEventDestroy(e);
    for (auto e : starPool) (void)hipEventDestroy(e);
    for (auto e : star2Pool) (void)hipEventDestroy(e);
    if (hostMaxDt) (void)hipHostFree(hostMaxDt);
    if (hostSc) (void)hipHostFree(hostSc);
    if (hostStat) (void)hipHostFree(hostStat);
    if (mlpDev) (void)hipFree(mlpDev);
    for (void *p : {(void *)velIn, (void *)velOut, (void *)dlist, (void *)ndef, (void *)sllfIn, (void *)sllfOut})
      if (p) (void)hipFree(p);
  }
};

namespace lsg {

static int prepareWidth(int spatial) {
  switch (spatial) {
  case 0: return 3;   // EO1 (lsEngquistOsher.hpp:31)
  case 1: return 5;   // EO2
  case 2: return 3;   // LF1 (lsLaxFriedrichs.hpp:35)
  case 3: return 5;   // LF2
  case 5: return 3;   // LLLF1 (lsLocalLocalLaxFriedrichs.hpp:32)
  case 6: return 5;   // LLLF2
  case 4: return 7;   // LLFA1 (lsLocalLaxFriedrichsAnalytical.hpp:51-55)
  case 7: return 7;   // LLF1 (lsLocalLaxFriedrichs.hpp:55)
  case 8: return 9;   // LLF2
  case 9: return 8;   // SLLF1 (lsStencilLocalLaxFriedrichsScalar.hpp:226, 2*(order+1)+4)
  case 10: return 5;  // WENO3 (lsWENO.hpp:44)
  case 11: return 7;  // WENO5
  case 12: return 7;  // WENO5 with learned weights (same stencil)
  default: return -1;
  }
}
static int schemeOrder(int spatial) {
  switch (spatial) {
  case 1: case 3: case 6: case 8: return 2;
  case 10: return 3;
  case 11: case 12: return 5;
  default: return 1;
  }
}

static int ensureRates(lsg_advect *a, size_t cap) {
  if (a->ratesCap < cap) {
    if (a->rates) (void)hipFree(a->rates);
    if (a->diss) (void)hipFree(a->diss);
    a->diss = nullptr;
    LSG_HIP(hipMalloc(&a->rates, cap * sizeof(double)));
    a->ratesCap = cap;
  }
  if (a->saveVelocities && !a->diss) LSG_HIP(hipMalloc(&a->diss, a->ratesCap * sizeof(double)));
  return 0;
}

// Point data after updateLevelSet: its Reduce(1) keeps the active points,
// each recording itself (lsReduce.hpp:105-111); with saveAdvectionVelocities
// the arrays "AdvectionVelocities" (the last rate = time * velocity) and
// "Dissipation" are inserted or replaced (lsAdvect.hpp:762-796).
static void pdUpdate(lsg_advect *a, lsg_domain *d, const double *dt, int nmat) {
  if (d->pd.empty() && !a->saveVelocities) return;
  if (pdSources(d, PdSrcParams{SRC_CENTRE, 0, 0}) != 0) return;
  if (pdApply(d) != 0) return;
  if (!a->saveVelocities) return;
  for (const char *l : {"AdvectionVelocities", "Dissipation"})  // velocityLabel, dissipationLabel (:833-834)
    if (!pdFind(d, l, 1)) {
      d->pd.emplace_back();
      d->pd.back().label = l;
    }
  if (pdEnsure(d, d->cur.b.cap) != 0) return;
  launchPdVelocities(g_ctx.s, g_pd->entry, g_pd->nPush, d->tmp.b.val, a->rates, a->diss, dt, nmat, d->cur.b.cap,
                     pdFind(d, "AdvectionVelocities", 1)->buf, pdFind(d, "Dissipation", 1)->buf);
}

// One forward-Euler stage (lsAdvectIntegrationSchemes.hpp:46-59):
// computeRates (prepareLS + integrateTime), updateLevelSet, rebuildLS.
// Host-evaluated velocities of the active points of the expanded band: the
// inputs of getScalarVelocity / getVectorVelocity are gathered on the device,
// the callbacks run on the host in point order, the results go back as the
// rate kernel's per-point velocities.
// With several level sets the callbacks run for every material m of every
// active point (velocities per (point, m), 4 doubles each): the material a
// point's scheme call uses is only known on the device.
// The LF schemes get findGlobalAlphas' call as a second record per point
// and the LLF schemes the 3^D alpha box (kVelInputs): the global alphas are
// then reduced here from the callbacks' results -- the maximum does not
// depend on the order -- and written to the device scalars.  The box and
// global-alpha calls are made for every active point, also where the
// reference skips the box (a zero Hamiltonian, lsLocalLaxFriedrichs.hpp:
// 178-180): the velocities used are the same, a callback counting its calls
// sees more of them.
static int hostVelocities(lsg_advect *a, lsg_domain *d, int nmat = 1, const LowerViews *lower = nullptr) {
  const size_t cap = d->cur.b.cap;
  const int R = hostVelocityRecords(a->spatial, d->D);
  if (a->velCap < cap || a->velMat < nmat || a->velRec < R) {
    for (void *p : {(void *)a->velIn, (void *)a->velOut, (void *)a->dlist}) if (p) (void)hipFree(p);
    a->velIn = a->velOut = nullptr;
    a->dlist = nullptr;
    a->velCap = 0;
    LSG_HIP(hipMalloc(&a->velIn, cap * (size_t)R * 8 * sizeof(double)));
    LSG_HIP(hipMalloc(&a->velOut, cap * (size_t)R * 4 * (size_t)nmat * sizeof(double)));
    LSG_HIP(hipMalloc(&a->dlist, cap * sizeof(uint32_t)));
    a->velCap = cap;
    a->velMat = nmat;
    a->velRec = R;
  }
  if (!a->ndef) LSG_HIP(hipMalloc(&a->ndef, sizeof(uint32_t)));
  LowerViews none{};
  launchVelInputs(g_ctx.s, d->D, d->cur.b, d->f, d->g, a->rp, lower ? *lower : none, g_ws->w, a->dlist, a->ndef,
                  a->velIn, g_ctx.sc);
  uint32_t nact = 0;
  LSG_HIP(hipMemcpyAsync(&nact, g_ws->w.ncand, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  const size_t nrec = (size_t)nact * R;
  a->velInH.resize(nrec * 8);
  a->velOutH.resize(nrec * 4 * nmat);
  if (nact) LSG_HIP(hipMemcpy(a->velInH.data(), a->velIn, nrec * 8 * sizeof(double), hipMemcpyDeviceToHost));
  // the analytical LLF's box records take getDissipationAlpha per direction
  // (its central differences in the normal slots), the others the velocities
  const bool alphaBox = a->spatial == 4;
  for (size_t r = 0; r < nrec; ++r) {
    const double *in = &a->velInH[r * 8];
    const uint64_t pid = (uint64_t)in[6];
    for (int m = 0; m < nmat; ++m) {
      double *o = &a->velOutH[(r * nmat + m) * 4];
      o[0] = o[1] = o[2] = o[3] = 0.0;
      if (alphaBox && r % (size_t)R != 0) {
        if (a->hostAlpha)
          for (int dir = 0; dir < d->D; ++dir) o[1 + dir] = a->hostAlpha(dir, m, in + 3, a->hostAlphaUser);
        continue;
      }
      o[0] = a->hostScalar ? a->hostScalar(in, m, in + 3, pid, a->hostUser) : 0.0;
      if (a->hostVector) a->hostVector(in, m, in + 3, pid, a->hostUser, o + 1);
    }
  }
  if (nact) LSG_HIP(hipMemcpy(a->velOut, a->velOutH.data(), nrec * 4 * nmat * sizeof(double), hipMemcpyHostToDevice));
  if (a->spatial == 9 && nact) {
    // stencil local Lax-Friedrichs: getScalarVelocity at the box normals
    // moved by -DN and +DN per axis, DN from the first round's box velocity
    // (lsStencilLocalLaxFriedrichsScalar.hpp:300-338); records without a
    // call in the reference (material -2) give 0
    const size_t nb = (size_t)nact * nmat * (R - 1) * 2 * d->D;
    if (a->sllfCap < nb) {
      for (void *p : {(void *)a->sllfIn, (void *)a->sllfOut}) if (p) (void)hipFree(p);
      a->sllfIn = a->sllfOut = nullptr;
      a->sllfCap = 0;
      LSG_HIP(hipMalloc(&a->sllfIn, nb * 8 * sizeof(double)));
      LSG_HIP(hipMalloc(&a->sllfOut, nb * sizeof(double)));
      a->sllfCap = nb;
    }
    launchSllfInputs(g_ctx.s, d->D, g_ws->w.ncand, nmat, a->velIn, a->velOut, a->rp.sllfBaseEps * d->g.dx,
                     a->sllfIn);
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    a->sllfInH.resize(nb * 8);
    a->sllfOutH.resize(nb);
    LSG_HIP(hipMemcpy(a->sllfInH.data(), a->sllfIn, nb * 8 * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < nb; ++r) {
      const double *in = &a->sllfInH[r * 8];
      const int m = (int)in[7];
      a->sllfOutH[r] = (m >= 0 && a->hostScalar) ? a->hostScalar(in, m, in + 3, (uint64_t)in[6], a->hostUser) : 0.0;
    }
    LSG_HIP(hipMemcpy(a->sllfOut, a->sllfOutH.data(), nb * sizeof(double), hipMemcpyHostToDevice));
  }
  if (R == 2) {  // LF: findGlobalAlphas from the second record of every point (lsAdvect.hpp:164-171)
    double al[3] = {0, 0, 0};
    for (uint32_t j = 0; j < nact; ++j) {
      const double *in = &a->velInH[((size_t)j * R + 1) * 8];
      const int m = std::max(0, std::min(nmat - 1, (int)in[7]));
      const double *o = &a->velOutH[(((size_t)j * R + 1) * nmat + m) * 4];
      for (int k = 0; k < d->D; ++k) {
        const double t = std::fabs((o[0] + o[1 + k]) * in[3 + k]);
        al[k] = (al[k] < t) ? t : al[k];  // std::max
      }
    }
    unsigned long long bits[3];
    std::memcpy(bits, al, sizeof(bits));
    LSG_HIP(hipMemcpyAsync(g_ctx.sc->alphaBits, bits, sizeof(bits), hipMemcpyHostToDevice, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
  }
  return 0;
}

static void stageFE(lsg_advect *a, lsg_domain *d, const double *maxDt, double *dtOut) {
  doExpand(d, prepareWidth(a->spatial));  // Advect::prepareLS (lsAdvect.hpp:973-1015)
  RateParams rp = a->rp;
  rp.hostVel = nullptr;
  rp.hostSllf = nullptr;
  rp.hostRec = hostVelocityRecords(a->spatial, d->D);
  rp.dissOut = a->saveVelocities ? a->diss : nullptr;
  const bool host = a->hostScalar || a->hostVector;
  if (host) {
    if (!a->hostErr) a->hostErr = hostVelocities(a, d);
    rp.hostVel = a->velOut;
    rp.hostSllf = a->sllfOut;
  }
  // LF global alphas: from the host calls, or the built-in functor here
  if ((a->spatial == 2 || a->spatial == 3) && !host)
    launchGlobalAlphas(g_ctx.s, d->D, d->cur.b, d->f, d->g, rp, g_ctx.sc);
  if (a->ignoreVoids && !a->hostErr) a->hostErr = markVoids(d, rp);
  if (a->profiling && a->evUsed + 2 <= (int)a->evPool.size()) {
    // events bracket the rate kernel itself
    hipEvent_t e0 = a->evPool[a->evUsed], e1 = a->evPool[a->evUsed + 1];
    a->evUsed += 2;
    launchRates(g_ctx.s, d->D, d->cur.b, d->f, d->g, rp, a->rates, g_ctx.sc, maxDt, dtOut, g_ws->w, e0, e1,
                a->capturing);
  } else {
    launchRates(g_ctx.s, d->D, d->cur.b, d->f, d->g, rp, a->rates, g_ctx.sc, maxDt, dtOut, g_ws->w);
  }
  // updateLevelSet: Reduce(1) + update (lsAdvect.hpp:686-800), then
  // rebuildLS; one star operation when no point data follows the band
  OpParams op{};
  op.rates = a->rates;
  op.dt = dtOut;
  if (fuseUpdateRebuild() && !singlePassBandOps() && d->pd.empty() && !a->saveVelocities) {
    op.kind = OP_UPDATE_REBUILD;
    op.a = (a->spatial == 9) ? 1.5 : 1.0;  // doRebuild's cutoff
    launchUpdRebuildOp(g_ctx.s, d->D, d->cur.b, d->tmp.b, d->f, d->g, g_ws->w, op, g_ctx.sc);
    swapCur(d);
    d->width = (a->spatial == 9) ? 3 : 2;
    return;
  }
  op.kind = OP_UPDATE;
  opSelf(d, op);
  pdUpdate(a, d, dtOut, 1);
  d->width = 1;
  doRebuild(d, a->spatial == 9, a->updatePointData);
}

// Advect::combineLevelSets (lsAdvect.hpp:199-236)
static void combine(lsg_advect *a, lsg_domain *d, lsg_domain *init, double wTarget, double wSource) {
  int steps = 1;
  if (a->temporal == 1) steps = 2;
  else if (a->temporal == 2) steps = 3;
  int ew = (int)std::ceil(2.0 * steps * a->cfl + 1);
  doExpand(d, ew);
  doExpand(init, ew);
  doBoolean(d, init, B_COMBINE, wTarget, wSource);
  doRebuild(d, a->spatial == 9, a->updatePointData);
}

// Copy without allocation (dst storage must already fit): usable in capture.
// (point data is copied too; it allocates, which only happens on steps
// launched directly: point data turns graph capture off)
static void copyDomainNoAlloc(lsg_domain *dst, const lsg_domain *src) {
  dst->D = src->D;
  dst->g = src->g;
  dst->f = src->f;
  dst->framed = src->framed;
  dst->width = src->width;
  launchCopyBand(g_ctx.s, src->cur.b, dst->cur.b, src->f);
  dst->hostN = src->hostN;
  (void)pdCopy(dst, src);
}

static int copyDomain(lsg_domain *dst, const lsg_domain *src) {
  dst->D = src->D;
  dst->g = src->g;
  for (int k = 0; k < 3; ++k) {
    dst->umin[k] = src->umin[k];
    dst->umax[k] = src->umax[k];
  }
  dst->f = src->f;
  dst->framed = src->framed;
  dst->width = src->width;
  if (int e = ensureCur(dst, src->cur.nwAlloc, src->cur.b.cap)) return e;
  launchCopyBand(g_ctx.s, src->cur.b, dst->cur.b, src->f);
  dst->hostN = src->hostN;
  return pdCopy(dst, src);
}

// Set while a velocity-update callback runs: advection calls made from it
// would overwrite the step's device scalars and are refused.
static bool g_inVelocityUpdate = false;

// The velocity-update callback between two stages of an RK step
// (lsAdvectIntegrationSchemes.hpp:87-89,128-130,148-150).  The stages so far
// are finished first; a step that already overflowed its band buffers
// returns 1 without calling back (the caller restarts it, so a restarted
// step can call back again).  The callback may read the level sets and
// change the velocity field; a change of the advected band is refused.
static int velocityUpdate(lsg_advect *a, lsg_domain *top) {
  LSG_HIP(hipMemcpyAsync(a->hostSc, g_ctx.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  if (a->hostSc->overflow) return 1;
  if (a->hostErr) return a->hostErr;
  const BandBufs b = top->cur.b;
  const Frame f = top->f;
  const int width = top->width;
  g_inVelocityUpdate = true;
  // the reference ignores the returned bool (0 or 1); a negative value is a
  // callback that failed (an exception in a language binding), which ends
  // the apply here as the exception ends the reference's
  const int r = a->velUpdate(top, a->velUpdateUser);
  g_inVelocityUpdate = false;
  if (r < 0) return fail(-46, "the velocity update callback failed; apply stopped at this stage");
  if (std::memcmp(&b, &top->cur.b, sizeof(BandBufs)) || std::memcmp(&f, &top->f, sizeof(Frame)) ||
      width != top->width)
    return fail(-1, "the velocity update callback must not modify the advected level set");
  return 0;
}

// One time step on the device (Advect::advect, lsAdvect.hpp:817-830).
// Returns 1 when a velocity-update callback found the step overflowed.
static int enqueueStep(lsg_advect *a, lsg_domain *d) {
  DevScalars *sc = g_ctx.sc;
  if (a->temporal == 0) {
    stageFE(a, d, &sc->maxTimeStep, &sc->dt[0]);
  } else if (a->temporal == 1) {  // evolveRungeKutta2 (lsAdvectIntegrationSchemes.hpp:61-108)
    copyDomainNoAlloc(a->init, d);
    stageFE(a, d, &sc->maxTimeStep, &sc->dt[0]);
    if (a->velUpdate)
      if (int r = velocityUpdate(a, d)) return r;
    stageFE(a, d, &sc->dt[0], &sc->dt[1]);
    combine(a, d, a->init, 0.5, 0.5);
  } else {  // evolveRungeKutta3 (lsAdvectIntegrationSchemes.hpp:110-168)
    copyDomainNoAlloc(a->init, d);
    stageFE(a, d, &sc->maxTimeStep, &sc->dt[0]);
    if (a->velUpdate)
      if (int r = velocityUpdate(a, d)) return r;
    stageFE(a, d, &sc->dt[0], &sc->dt[1]);
    combine(a, d, a->init, 0.75, 0.25);
    if (a->velUpdate)
      if (int r = velocityUpdate(a, d)) return r;
    stageFE(a, d, &sc->dt[0], &sc->dt[2]);
    combine(a, d, a->init, 1.0 / 3.0, 2.0 / 3.0);
  }
  launchStepTime(g_ctx.s, sc, a->temporal);
  return 0;
}

struct StepOut {
  double stepTime, dt1;
  unsigned long long active;
};

static uint64_t fnv(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = (const unsigned char *)p;
  for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
  return h;
}

// Everything a captured step bakes into its kernel arguments.
static uint64_t stepKey(lsg_advect *a, lsg_domain *d) {
  uint64_t h = 1469598103934665603ull;
  auto add = [&](const void *p, size_t n) { h = fnv(h, p, n); };
  add(&d->cur.b, sizeof(BandBufs));
  add(&d->tmp.b, sizeof(BandBufs));
  add(&d->f, sizeof(Frame));
  add(&d->width, sizeof(int));
  if (a->init) {
    add(&a->init->cur.b, sizeof(BandBufs));
    add(&a->init->tmp.b, sizeof(BandBufs));
  }
  add(&a->save->cur.b, sizeof(BandBufs));
  add(&g_ws->w, sizeof(Workspace));
  add(&a->rp, sizeof(RateParams));
  add(&a->temporal, sizeof(int));
  add(&a->cfl, sizeof(double));
  add(&a->rates, sizeof(double *));
  add(&a->profiling, sizeof(bool));
  const bool singlePass = singlePassBandOps();  // the band-operation kernels a capture records
  add(&singlePass, sizeof(bool));
  return h;
}

// The device work of one step: prologue (reset flags, take maxTimeStep,
// keep a copy for retries), the stages, the returned time, band statistics.
static int enqueueStepBody(lsg_advect *a, lsg_domain *d) {
  launchStepBegin(g_ctx.s, g_ctx.sc, a->devMaxDt);
  copyDomainNoAlloc(a->save, d);
  a->evUsed = 0;
  a->starT = StarTiming{a->starPool.data(), (int)a->starPool.size(), 0, a->capturing};
  a->star2T = StarTiming{a->star2Pool.data(), (int)a->star2Pool.size(), 0, a->capturing};
  if (a->profiling) {
    g_starTiming = &a->starT;
    g_star2Timing = &a->star2T;
  }
  const int r = enqueueStep(a, d);
  g_starTiming = nullptr;
  g_star2Timing = nullptr;
  if (r) return r;
  launchFrameCheck(g_ctx.s, d->cur.b, d->f, d->g, d->D, REFRAME_SLACK, g_ctx.sc);
  return 0;
}

static int runStep(lsg_advect *a, lsg_domain *d, double maxTimeStep, StepOut &out) {
  for (int attempt = 0; attempt < 6; ++attempt) {
    // allocation happens here, never inside a capture
    if (!a->save) a->save = new lsg_domain();
    if (int e = ensureCur(a->save, d->cur.nwAlloc, d->cur.b.cap)) return e;
    if (a->init) {
      if (int e = ensureCur(a->init, d->cur.nwAlloc, d->cur.b.cap)) return e;
      a->init->f = d->f;
      if (int e = ensureStores(a->init, d->cur.b.cap)) return e;
    }
    if (int e = ensureStores(d, d->cur.b.cap)) return e;
    int widthBefore = d->width;
    *a->hostMaxDt = maxTimeStep;
    bool launched = false;
    const bool hostVel = a->hostScalar || a->hostVector;
    a->hostErr = 0;
    // point data allocates during the step: such steps are launched directly
    const bool pointData = !d->pd.empty() || a->saveVelocities || a->ignoreVoids;
    if (!pointData) {
      pdClear(a->save);
      if (a->init) pdClear(a->init);
    }
    // a velocity-update callback runs between the stages: no graph
    const bool callback = a->velUpdate && a->temporal != 0;
    if (a->useGraphs && !hostVel && !pointData && !callback) {
      uint64_t key = stepKey(a, d);
      lsg_advect::StepGraph *g = nullptr;
      for (auto &c : a->graphs)
        if (c.exec && c.key == key) g = &c;
      if (!g) {
        g = &a->graphs[a->nextGraph];
        a->nextGraph ^= 1;
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        *g = lsg_advect::StepGraph{};
        hipGraph_t graph = nullptr;
        bool ok = hipStreamBeginCapture(g_ctx.s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
          a->capturing = true;
          (void)enqueueStepBody(a, d);  // no callback while capturing
          a->capturing = false;
          ok = hipStreamEndCapture(g_ctx.s, &graph) == hipSuccess && graph;
        }
        if (ok) ok = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
          // capture unsupported here: fall back to direct launches for good
          (void)hipGetLastError();
          a->useGraphs = false;
          a->dropGraphs();
          return fail(-31, "HIP graph capture failed; retry the call (direct launches are used from now on)");
        }
        g->key = key;
        g->dCur = d->cur;
        g->dTmp = d->tmp;
        if (a->init) {
          g->iCur = a->init->cur;
          g->iTmp = a->init->tmp;
          g->iWidth = a->init->width;
        }
        g->dWidth = d->width;
        g->evUsed = a->evUsed;
        g->starUsed = a->starT.used;
      } else {
        // replay: the host bookkeeping takes the state the captured step leaves
        d->cur = g->dCur;
        d->tmp = g->dTmp;
        d->width = g->dWidth;
        if (a->init) {
          a->init->cur = g->iCur;
          a->init->tmp = g->iTmp;
          a->init->width = g->iWidth;
          a->init->f = d->f;
          a->init->framed = true;
        }
        a->evUsed = g->evUsed;
        a->starT.used = g->starUsed;
      }
      LSG_HIP(hipGraphLaunch(g->exec, g_ctx.s));
      launched = true;
    }
    if (!launched)
      if (int e = enqueueStepBody(a, d); e < 0) return e;
    if (a->hostErr) return a->hostErr;
    LSG_HIP(hipMemcpyAsync(a->hostSc, g_ctx.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    const DevScalars &h = *a->hostSc;
    if (lookbackFailed(h)) {  // repeat the step from the saved band
      a->dropGraphs();
      if (int e = copyDomain(d, a->save)) return e;
      d->width = widthBefore;
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      continue;
    }
    if (!h.overflow) {
      out.stepTime = h.stepTime;
      out.dt1 = h.dt[0];
      out.active = h.totalActive - h.stepActive0;
      if (a->profiling) {
        for (int k = 0; k + 1 < a->evUsed; k += 2) {
          float ms = 0;
          if (hipEventElapsedTime(&ms, a->evPool[k], a->evPool[k + 1]) == hipSuccess) {
            a->rateMs += ms;
            a->rateLaunches += 1;
          }
        }
        for (int k = 0; k + 1 < a->starT.used; k += 2) {
          float ms = 0;
          if (hipEventElapsedTime(&ms, a->starPool[k], a->starPool[k + 1]) == hipSuccess) {
            a->starMs += ms;
            a->starLaunches += 1;
          }
        }
        for (int k = 0; k + 1 < a->star2T.used; k += 2) {
          float ms = 0;
          if (hipEventElapsedTime(&ms, a->star2Pool[k], a->star2Pool[k + 1]) == hipSuccess) {
            a->star2Ms += ms;
            a->star2Launches += 1;
          }
        }
      }
      // re-frame when the band approached the frame edge in this step
      if (h.reframe) {
        a->dropGraphs();
        return maybeReframe(d, true);
      }
      return 0;
    }
    // grow and retry from the saved band
    a->dropGraphs();
    size_t cap = d->cur.b.cap * 2;
    if (int e = copyDomain(d, a->save)) return e;
    d->width = widthBefore;
    if (int e = ensureStores(d, cap)) return e;
    if (a->init && a->init->framed && (a->init->cur.b.cap < cap))
      if (int e = ensureStores(a->init, cap)) return e;
    if (int e = ensureRates(a, cap)) return e;
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
  }
  return fail(-30, "band capacity overflow persists after retries");
}

// ---------------------------------------------------- several level sets ---
// Advect with level sets below the advected top (lsAdvect.hpp:435-612,
// 686-813; lsAdvectIntegrationSchemes.hpp:46-168).  All level sets share
// the top's frame during a step.  Steps are launched directly (no graph): a
// combine's movedDown flag is read on the host to decide whether the lower
// level sets are restored (RK3's last combine, and with a velocity-update
// callback every combine).

static void adjustLowerLayers(lsg_advect *a) {  // lsAdvect.hpp:802-813
  if (a->spatial == 9) return;  // the stencil local Lax-Friedrichs scheme keeps the lower level sets
  lsg_domain *top = a->levelSets.back();
  for (size_t i = 0; i + 1 < a->levelSets.size(); ++i) doBoolean(a->levelSets[i], top, B_INTERSECT, 0, 0);
}

static void stageMM(lsg_advect *a, const double *maxDt, double *dtOut) {
  lsg_domain *top = a->levelSets.back();
  const int L = (int)a->levelSets.size();
  doExpand(top, prepareWidth(a->spatial));
  LowerViews lw{};
  lw.n = L - 1;
  for (int m = 0; m < L - 1; ++m) {
    lw.bits[m] = a->levelSets[m]->cur.b.bits;
    lw.pre[m] = a->levelSets[m]->cur.b.pre;
    lw.val[m] = a->levelSets[m]->cur.b.val;
  }
  RateParams rp = a->rp;
  rp.dissOut = a->saveVelocities ? a->diss : nullptr;
  rp.hostVel = nullptr;
  rp.hostSllf = nullptr;
  rp.hostRec = hostVelocityRecords(a->spatial, top->D);
  const bool host = a->hostScalar || a->hostVector;
  if (host) {
    if (!a->hostErr) a->hostErr = hostVelocities(a, top, L, &lw);
    rp.hostVel = a->velOut;
    rp.hostSllf = a->sllfOut;
  }
  if ((a->spatial == 2 || a->spatial == 3) && !host)
    launchGlobalAlphasMM(g_ctx.s, top->D, top->cur.b, top->f, top->g, lw, rp, g_ctx.sc);
  if (a->ignoreVoids && !a->hostErr) a->hostErr = markVoids(top, rp);
  launchRatesMM(g_ctx.s, top->D, top->cur.b, top->f, top->g, lw, rp, a->rates, g_ctx.sc, maxDt, dtOut, g_ws->w);
  OpParams op{};
  op.kind = OP_UPDATE_MM;
  op.rates = a->rates;
  op.dt = dtOut;
  op.nmat = L;
  opSelf(top, op);
  pdUpdate(a, top, dtOut, L);
  top->width = 1;
  doRebuild(top, a->spatial == 9, a->updatePointData);
}

static void combineMM(lsg_advect *a, double wTarget, double wSource, uint32_t *movedDown) {
  lsg_domain *top = a->levelSets.back();
  const int ew = (int)std::ceil(2.0 * (a->temporal == 1 ? 2 : 3) * a->cfl + 1);
  doExpand(top, ew);
  doExpand(a->init, ew);
  doBoolean(top, a->init, B_COMBINE, wTarget, wSource, movedDown);
  doRebuild(top, a->spatial == 9, a->updatePointData);
}

static int enqueueStepMM(lsg_advect *a) {
  DevScalars *sc = g_ctx.sc;
  lsg_domain *top = a->levelSets.back();
  const size_t nl = a->levelSets.size() - 1;
  launchStepBegin(g_ctx.s, sc, a->devMaxDt);
  copyDomainNoAlloc(a->save, top);
  for (size_t i = 0; i < nl; ++i) copyDomainNoAlloc(a->lowerSave[i], a->levelSets[i]);
  const bool cb = a->velUpdate != nullptr;
  // a combine that moved the surface down anywhere sends the lower level sets
  // back to the start of the step (lsAdvectIntegrationSchemes.hpp:98-103,
  // 139-144,157-163)
  auto combineRestore = [&](double wTarget, double wSource) -> int {
    launchSetFlag(g_ctx.s, &sc->movedDown, 0);
    combineMM(a, wTarget, wSource, &sc->movedDown);
    LSG_HIP(hipMemcpyAsync(a->hostSc, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    if (a->hostSc->movedDown && !a->hostSc->overflow)
      for (size_t i = 0; i < nl; ++i) copyDomainNoAlloc(a->levelSets[i], a->lowerSave[i]);
    return 0;
  };
  if (a->temporal == 0) {
    stageMM(a, &sc->maxTimeStep, &sc->dt[0]);
    adjustLowerLayers(a);
  } else if (a->temporal == 1) {
    copyDomainNoAlloc(a->init, top);
    stageMM(a, &sc->maxTimeStep, &sc->dt[0]);
    if (cb) {  // with a callback the first stage updates the lower level sets
      adjustLowerLayers(a);
      if (int r = velocityUpdate(a, top)) return r;
    }
    stageMM(a, &sc->dt[0], &sc->dt[1]);
    if (cb) {
      if (int r = combineRestore(0.5, 0.5)) return r;
    } else {
      combineMM(a, 0.5, 0.5, nullptr);
    }
    adjustLowerLayers(a);
  } else {
    copyDomainNoAlloc(a->init, top);
    stageMM(a, &sc->maxTimeStep, &sc->dt[0]);
    if (cb) {
      adjustLowerLayers(a);
      if (int r = velocityUpdate(a, top)) return r;
    }
    stageMM(a, &sc->dt[0], &sc->dt[1]);
    if (cb) {
      if (int r = combineRestore(0.75, 0.25)) return r;
    } else {
      combineMM(a, 0.75, 0.25, nullptr);
    }
    adjustLowerLayers(a);
    if (cb)
      if (int r = velocityUpdate(a, top)) return r;
    stageMM(a, &sc->dt[0], &sc->dt[2]);
    if (int r = combineRestore(1.0 / 3.0, 2.0 / 3.0)) return r;
    adjustLowerLayers(a);
  }
  launchStepTime(g_ctx.s, sc, a->temporal);
  launchFrameCheck(g_ctx.s, top->cur.b, top->f, top->g, top->D, REFRAME_SLACK, sc);
  return 0;
}

static int runStepMM(lsg_advect *a, double maxTimeStep, StepOut &out) {
  lsg_domain *top = a->levelSets.back();
  const size_t L = a->levelSets.size();
  for (int attempt = 0; attempt < 6; ++attempt) {
    if (int e = commonFrameAll(a->levelSets)) return e;
    size_t cap = top->cur.b.cap;
    for (lsg_domain *d : a->levelSets) cap = std::max(cap, d->cur.b.cap);
    for (lsg_domain *d : a->levelSets)
      if (int e = ensureStores(d, cap)) return e;
    if (!a->save) a->save = new lsg_domain();
    if (int e = ensureCur(a->save, top->cur.nwAlloc, cap)) return e;
    while (a->lowerSave.size() < L - 1) a->lowerSave.push_back(new lsg_domain());
    for (size_t i = 0; i + 1 < L; ++i)
      if (int e = ensureCur(a->lowerSave[i], top->cur.nwAlloc, cap)) return e;
    if (a->init) {
      if (int e = ensureCur(a->init, top->cur.nwAlloc, cap)) return e;
      a->init->f = top->f;
      if (int e = ensureStores(a->init, cap)) return e;
    }
    if (int e = ensureRates(a, cap * 2 * L)) return e;
    std::vector<int> widths;
    for (lsg_domain *d : a->levelSets) widths.push_back(d->width);
    *a->hostMaxDt = maxTimeStep;
    a->hostErr = 0;
    if (int e = enqueueStepMM(a); e < 0) return e;  // 1: overflow seen before a callback
    if (a->hostErr) return a->hostErr;
    LSG_HIP(hipMemcpyAsync(a->hostSc, g_ctx.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    const DevScalars &h = *a->hostSc;
    if (lookbackFailed(h)) {  // repeat the step from the saved bands
      if (int e = copyDomain(top, a->save)) return e;
      for (size_t i = 0; i + 1 < L; ++i)
        if (int e = copyDomain(a->levelSets[i], a->lowerSave[i])) return e;
      for (size_t i = 0; i < L; ++i) a->levelSets[i]->width = widths[i];
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      continue;
    }
    if (!h.overflow) {
      out.stepTime = h.stepTime;
      out.dt1 = h.dt[0];
      out.active = h.totalActive - h.stepActive0;
      if (h.reframe) return maybeReframe(top, true);  // the next step re-frames the others
      return 0;
    }
    // grow every band and retry from the start of the step
    if (int e = copyDomain(top, a->save)) return e;
    for (size_t i = 0; i + 1 < L; ++i)
      if (int e = copyDomain(a->levelSets[i], a->lowerSave[i])) return e;
    for (size_t i = 0; i < L; ++i) a->levelSets[i]->width = widths[i];
    for (lsg_domain *d : a->levelSets)
      if (int e = ensureStores(d, cap * 2)) return e;
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
  }
  return fail(-30, "band capacity overflow persists after retries");
}

static int advectApply(lsg_advect *a) {
  if (a->levelSets.empty()) return fail(-1, "No level sets passed to Advect. Not advecting.");
  if (!a->velSet) return fail(-2, "No velocity field passed to Advect. Not advecting.");
  const bool multi = a->levelSets.size() > 1;
  if (multi) {
    if (a->levelSets.size() > (size_t)MAX_LEVEL_SETS) return fail(-3, "Advect: at most 8 level sets are supported");
    if (a->spatial == 12) return fail(-3, "Advect: with several level sets the learned WENO5 scheme is not supported");
  }
  if (prepareWidth(a->spatial) < 0) return fail(-4, "Advect: Discretization scheme not found.");
  if (!(a->cfl > 0)) return fail(-6, "Advect: time step ratio must be positive");
  if (int e = ctxInit()) return e;
  lsg_domain *d = a->levelSets.back();
  if (!d->framed) return fail(-7, "Advect: level set has no data");
  // scan grid by band size for this apply (LSG_SCAN_BLOCKS overrides it):
  // the default up to band stores of 8M entries (about 1M band entries),
  // the largest grid beyond (rough1024 LLF2: 11.47 ms per step at 4096
  // blocks, 11.70 at 3072, 12.17 at 2304, 12.39 at 1536; trench256 WENO5
  // is fastest at 1536, 1.44 ms against 1.72 ms at 3072).  Every step graph
  // keeps the grid it was captured with, so its kernels agree on the
  // per-block counts.
  struct ScanGridScope {
    ~ScanGridScope() { setScanBlocks(defaultScanBlocks()); }
  } scanGridScope;
  if (!getenv("LSG_SCAN_BLOCKS"))
    setScanBlocks(d->cur.b.cap >= ((size_t)8 << 20) ? SCAN_BLOCKS_MAX : SCAN_BLOCKS);
  if (a->cfl >= 0.5) fprintf(stderr, "[viennals_gpu] Integration time step ratio should be smaller than 0.5. Advection might fail!\n");
  a->rp.spatial = a->spatial;
  a->rp.order = schemeOrder(a->spatial);
  a->rp.cfl = a->cfl;
  a->rp.alphaFactor = a->dissipationAlpha;
  a->rp.checkDissipation = a->checkDissipation;
  a->rp.sllfBaseEps = std::cbrt(std::numeric_limits<double>::epsilon());
  // WENO receives dissipationAlpha as its calcNormal flag (lsAdvect.hpp:670-676)
  a->rp.calcNormals = (a->spatial >= 10) ? (a->dissipationAlpha != 0.0) : a->calcNormals;
  if ((a->hostScalar || a->hostVector) && a->spatial == 12)
    return fail(-9, "Advect: host-evaluated velocities are not supported with the learned WENO5 scheme");
  if (a->spatial == 12 && !a->mlpDev)
    return fail(-8, "Advect: learned WENO5 selected but no network parameters were set");
  a->rp.mlp = a->mlpDev;
  // the network on the matrix cores per wave (default) or per lane
  // (LSG_LEARNED_WAVE=0)
  {
    const char *e = getenv("LSG_LEARNED_WAVE");
    a->rp.learnedWave = (e && e[0] == '0') ? 0 : 1;
  }
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  if (int e = ensureRates(a, d->cur.b.cap)) return e;
  if (a->temporal != 0) {
    if (!a->init) a->init = new lsg_domain();
    if (int e = copyDomain(a->init, d)) return e;
    if (int e = ensureStores(a->init, d->cur.b.cap)) return e;
  }
  if (!a->hostMaxDt) {
    LSG_HIP(hipHostMalloc((void **)&a->hostMaxDt, sizeof(double), hipHostMallocMapped | hipHostMallocCoherent));
    LSG_HIP(hipHostGetDevicePointer((void **)&a->devMaxDt, a->hostMaxDt, 0));
    LSG_HIP(hipHostMalloc((void **)&a->hostSc, sizeof(DevScalars), hipHostMallocDefault));
    LSG_HIP(hipHostMalloc((void **)&a->hostStat, 8 * sizeof(uint32_t), hipHostMallocDefault));
  }
  if (a->profiling && a->evPool.empty()) {
    a->evPool.resize(8);
    for (auto &e : a->evPool) LSG_HIP(hipEventCreate(&e));
    a->starPool.resize(96);
    for (auto &e : a->starPool) LSG_HIP(hipEventCreate(&e));
    a->star2Pool.resize(64);
    for (auto &e : a->star2Pool) LSG_HIP(hipEventCreate(&e));
  }
  if (const char *ng = getenv("LSG_NO_GRAPH")) a->useGraphs = (ng[0] == '0');
  a->pointUpdates = 0;
  StepOut so{};
  auto step = [&](double maxT) { return multi ? runStepMM(a, maxT, so) : runStep(a, d, maxT, so); };
  if (a->advectionTime == 0.) {
    if (int e = step(std::numeric_limits<double>::max())) return e;
    a->advectedTime = so.stepTime;
    a->numSteps = 1;
    a->pointUpdates += (double)so.active;
  } else {
    double t = 0.0;
    a->numSteps = 0;
    while (t < a->advectionTime) {
      if (int e = step(a->advectionTime - t)) return e;
      a->pointUpdates += (double)so.active;
      if (!(so.stepTime > 0.)) return fail(-8, "Advect: time step is not positive; advection stalled");
      t += so.stepTime;
      ++a->numSteps;
      if (a->singleStep) break;
    }
    a->advectedTime = t;
  }
  a->currentTimeStep = so.dt1;
  if (a->temporal == 0) a->currentTimeStep = so.stepTime;
  if (a->profiling) {
    a->rateBytes = a->hostSc->rateBytes;
    a->starBytes = a->hostSc->starBytes;
    a->star2Bytes = a->hostSc->star2Bytes;
  }
  return 0;
}


// ================================================================ batch ===
//
// A batch advects many independent level sets with one launch per kernel:
// their bands are concatenated into one segmented band (bitmap words, entries
// and candidates tagged with their level set), every level set keeps its own
// frame, grid, velocity and CFL time step (SegInfo table).  A step of the
// batch is one Advect::apply with advectionTime 0 for every level set, and
// is captured into a HIP graph like the single-level-set step.

struct BStore {
  BandBufs b{};
  int64_t nwAlloc = 0;
  void release() {
    Store s;
    s.b = b;
    s.nwAlloc = nwAlloc;
    s.release();
    if (b.seg) (void)hipFree(b.seg);
    b = BandBufs{};
    nwAlloc = 0;
  }
  int alloc(int64_t nw, size_t cap) {
    release();
    Store s;
    if (int e = s.alloc(nw, cap)) return e;
    b = s.b;
    s.b = BandBufs{};
    nwAlloc = nw;
    LSG_HIP(hipMalloc(&b.seg, cap * sizeof(uint16_t)));
    return 0;
  }
};

__global__ void kAddCopy32(const uint32_t *src, uint32_t *dst, uint32_t n, uint32_t add) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i] + add;
}
__global__ void kFill16(uint16_t *p, uint16_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void kSetScalar32(uint32_t *p, uint32_t v) { *p = v; }
__global__ void kZero64(uint64_t *p) { *p = 0; }
__global__ void kGather32(const uint32_t *src, const uint32_t *idx, uint32_t *dst, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
static int gridOf(size_t n) { return (int)std::min<size_t>(4096, std::max<size_t>(1, (n + 255) / 256)); }

}  // namespace lsg

struct lsg_batch {
  int D = 3;
  std::vector<lsg_domain *> doms;
  std::vector<lsg::SegInfo> segsH;
  std::vector<int> velSet;
  lsg::SegInfo *segsD = nullptr;
  uint16_t *wsegD = nullptr;
  int64_t W = 0;
  uint32_t totalBase = 0;
  bool assembled = false;
  lsg::BStore cur, tmp, icur, itmp, save;
  int width = 2, iwidth = 2;
  int spatial = 11, temporal = 2;
  double cfl = 0.4;
  lsg::RateParams rp{};
  double *rates = nullptr;
  size_t ratesCap = 0;
  unsigned long long *segMax = nullptr;
  unsigned long long *segAlpha = nullptr, *segGA = nullptr;  // Lax-Friedrichs alphas per level set (3 each)
  double *segMaxDt = nullptr, *segDt = nullptr, *segStepTime = nullptr, *segAdvTime = nullptr;
  uint32_t *bounds = nullptr, *boundsIdx = nullptr;  // pre[] at the level-set word boundaries (write-back)
  double pointUpdates = 0;
  unsigned steps = 0;
  int scanBlocks = lsg::defaultScanBlocks();
  bool useGraphs = true;
  struct Graph {
    uint64_t key = 0;
    hipGraphExec_t exec = nullptr;
    lsg::BStore cur, tmp, icur, itmp;
    int width = 0, iwidth = 0;
  };
  Graph graphs[2];
  int nextGraph = 0;
  lsg::DevScalars *hostSc = nullptr;
  void dropGraphs() {
    for (auto &g : graphs) {
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
      g = Graph{};
    }
  }
  void freeDevice() {
    dropGraphs();
    cur.release();
    tmp.release();
    icur.release();
    itmp.release();
    save.release();
    for (void *p : {(void *)segsD, (void *)wsegD, (void *)rates, (void *)segMax, (void *)segMaxDt, (void *)segDt,
                    (void *)segStepTime, (void *)bounds, (void *)boundsIdx, (void *)segAlpha, (void *)segGA})
      if (p) (void)hipFree(p);
    segsD = nullptr;
    wsegD = nullptr;
    rates = nullptr;
    ratesCap = 0;
    segMax = nullptr;
    segAlpha = segGA = nullptr;
    segMaxDt = segDt = segStepTime = nullptr;
    bounds = boundsIdx = nullptr;
  }
  ~lsg_batch() {
    freeDevice();
    if (segAdvTime) (void)hipFree(segAdvTime);
    if (hostSc) (void)hipHostFree(hostSc);
  }
};

namespace lsg {

static SegCtx segCtx(const lsg_batch *b) {
  return SegCtx{b->segsD, (uint32_t)b->doms.size(), b->totalBase, b->wsegD, (uint32_t)b->W};
}

static int ensureBatchWs(size_t cap, int64_t nw) {
  if (!g_ws) g_ws = new WsStore();
  if (g_ws->w.ccap < cap || g_ws->nwAlloc < nw || !g_ws->w.cseg) {
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    if (int e = g_ws->alloc(std::max(g_ws->nwAlloc, nw), std::max(g_ws->w.ccap, cap))) return e;
  }
  return 0;
}

// Concatenate the level sets into one segmented band.
static int batchAssemble(lsg_batch *b) {
  const size_t S = b->doms.size();
  b->freeDevice();
  std::vector<uint32_t> nEnt(S);
  b->segsH.resize(S);
  int64_t W = 0;
  uint32_t baseAcc = 0;
  uint64_t entAcc = 0;
  int width = 1 << 30;
  for (size_t k = 0; k < S; ++k) {
    lsg_domain *d = b->doms[k];
    LSG_HIP(hipMemcpyAsync(&nEnt[k], d->cur.b.n, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
    SegInfo &s = b->segsH[k];
    s.f = d->f;
    s.g = d->g;
    s.wordOff = (uint32_t)W;
    baseAcc += (uint32_t)d->f.base;
    s.baseIncl = baseAcc;
    W += d->f.nwords;
    if (k > 0 && d->width != width) return fail(-41, "batch: the level sets must have the same band width");
    if (!d->pd.empty()) return fail(-41, "batch: level sets with point data are not supported");
    width = d->width;
  }
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  for (size_t k = 0; k < S; ++k) entAcc += nEnt[k];
  if (W >= (int64_t)UINT32_MAX - 1 || entAcc >= (uint64_t)UINT32_MAX / 2)
    return fail(-40, "batch too large: more than 2^32 bitmap words or 2^31 entries");
  b->W = W;
  b->totalBase = baseAcc;
  b->width = width;
  const size_t cap = std::max<size_t>((size_t)1 << 20, (size_t)(entAcc * 3 / 2 + 4096));
  if (int e = b->cur.alloc(W, cap)) return e;
  if (int e = b->tmp.alloc(W, cap)) return e;
  if (int e = b->save.alloc(W, cap)) return e;
  if (b->temporal != 0) {
    if (int e = b->icur.alloc(W, cap)) return e;
    if (int e = b->itmp.alloc(W, cap)) return e;
  }
  LSG_HIP(hipMalloc(&b->segsD, S * sizeof(SegInfo)));
  LSG_HIP(hipMalloc(&b->wsegD, (W + 1) * sizeof(uint16_t)));
  LSG_HIP(hipMalloc(&b->rates, cap * sizeof(double)));
  b->ratesCap = cap;
  LSG_HIP(hipMalloc(&b->segMax, S * sizeof(unsigned long long)));
  LSG_HIP(hipMalloc(&b->segAlpha, 3 * S * sizeof(unsigned long long)));
  LSG_HIP(hipMalloc(&b->segGA, 3 * S * sizeof(unsigned long long)));
  LSG_HIP(hipMalloc(&b->segMaxDt, S * sizeof(double)));
  LSG_HIP(hipMalloc(&b->segDt, 3 * S * sizeof(double)));
  LSG_HIP(hipMalloc(&b->segStepTime, S * sizeof(double)));
  LSG_HIP(hipMalloc(&b->bounds, (S + 1) * sizeof(uint32_t)));
  LSG_HIP(hipMalloc(&b->boundsIdx, (S + 1) * sizeof(uint32_t)));
  if (!b->segAdvTime) {
    LSG_HIP(hipMalloc(&b->segAdvTime, S * sizeof(double)));
    LSG_HIP(hipMemset(b->segAdvTime, 0, S * sizeof(double)));
  }
  LSG_HIP(hipMemcpyAsync(b->segsD, b->segsH.data(), S * sizeof(SegInfo), hipMemcpyHostToDevice, g_ctx.s));
  // copy every level set's arrays into place
  uint32_t inFrameBefore = 0;
  uint64_t e0 = 0;
  BandBufs &o = b->cur.b;
  for (size_t k = 0; k < S; ++k) {
    lsg_domain *d = b->doms[k];
    const SegInfo &s = b->segsH[k];
    const int64_t nw = d->f.nwords;
    LSG_HIP(hipMemcpyAsync(o.bits + s.wordOff, d->cur.b.bits, nw * sizeof(uint64_t), hipMemcpyDeviceToDevice, g_ctx.s));
    kAddCopy32<<<gridOf(nw), 256, 0, g_ctx.s>>>(d->cur.b.pre, o.pre + s.wordOff, (uint32_t)nw, inFrameBefore);
    LSG_HIP(hipMemcpyAsync(o.val + e0, d->cur.b.val, nEnt[k] * sizeof(double), hipMemcpyDeviceToDevice, g_ctx.s));
    LSG_HIP(hipMemcpyAsync(o.pos + e0, d->cur.b.pos, nEnt[k] * sizeof(uint32_t), hipMemcpyDeviceToDevice, g_ctx.s));
    kFill16<<<gridOf(nEnt[k]), 256, 0, g_ctx.s>>>(o.seg + e0, (uint16_t)k, nEnt[k]);
    kFill16<<<gridOf(nw), 256, 0, g_ctx.s>>>(b->wsegD + s.wordOff, (uint16_t)k, (size_t)nw);
    inFrameBefore += nEnt[k] - (uint32_t)d->f.base;
    e0 += nEnt[k];
  }
  kZero64<<<1, 1, 0, g_ctx.s>>>(o.bits + W);
  kSetScalar32<<<1, 1, 0, g_ctx.s>>>(o.pre + W, inFrameBefore);
  kSetScalar32<<<1, 1, 0, g_ctx.s>>>(o.n, (uint32_t)e0);
  kFill16<<<1, 1, 0, g_ctx.s>>>(b->wsegD + W, (uint16_t)(S - 1), 1);
  std::vector<uint32_t> bi(S + 1);
  for (size_t k = 0; k < S; ++k) bi[k] = b->segsH[k].wordOff;
  bi[S] = (uint32_t)W;
  LSG_HIP(hipMemcpyAsync(b->boundsIdx, bi.data(), (S + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  if (int e = ensureBatchWs(cap, W)) return e;
  b->assembled = true;
  return 0;
}

// Copy the level sets back into their domains.
static int batchWriteBack(lsg_batch *b) {
  if (!b->assembled) return 0;
  const size_t S = b->doms.size();
  BandBufs &o = b->cur.b;
  kGather32<<<(int)((S + 1 + 255) / 256), 256, 0, g_ctx.s>>>(o.pre, b->boundsIdx, b->bounds, (uint32_t)(S + 1));
  std::vector<uint32_t> bd(S + 1);
  LSG_HIP(hipMemcpyAsync(bd.data(), b->bounds, (S + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  for (size_t k = 0; k < S; ++k) {
    lsg_domain *d = b->doms[k];
    const SegInfo &s = b->segsH[k];
    const uint32_t base = (uint32_t)s.f.base;
    const uint32_t n = bd[k + 1] - bd[k] + base;
    const uint64_t e0 = (uint64_t)bd[k] + s.baseIncl - base;
    d->f = s.f;
    d->framed = true;
    if (int e = ensureStores(d, std::max(capFor(n), d->cur.b.cap))) return e;
    const int64_t nw = s.f.nwords;
    LSG_HIP(hipMemcpyAsync(d->cur.b.bits, o.bits + s.wordOff, nw * sizeof(uint64_t), hipMemcpyDeviceToDevice, g_ctx.s));
    kZero64<<<1, 1, 0, g_ctx.s>>>(d->cur.b.bits + nw);
    kAddCopy32<<<gridOf(nw + 1), 256, 0, g_ctx.s>>>(o.pre + s.wordOff, d->cur.b.pre, (uint32_t)(nw + 1),
                                                     (uint32_t)(0u - bd[k]));
    LSG_HIP(hipMemcpyAsync(d->cur.b.val, o.val + e0, n * sizeof(double), hipMemcpyDeviceToDevice, g_ctx.s));
    LSG_HIP(hipMemcpyAsync(d->cur.b.pos, o.pos + e0, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, g_ctx.s));
    kSetScalar32<<<1, 1, 0, g_ctx.s>>>(d->cur.b.n, n);
    d->hostN = n;
    d->width = b->width;
  }
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

static void bSwap(lsg_batch *b) { std::swap(b->cur, b->tmp); }
static void bSwapInit(lsg_batch *b) { std::swap(b->icur, b->itmp); }

static void bStar(lsg_batch *b, bool init, const OpParams &op) {
  BStore &in = init ? b->icur : b->cur;
  BStore &out = init ? b->itmp : b->tmp;
  launchStarOpSeg(g_ctx.s, b->D, in.b, out.b, segCtx(b), g_ws->w, op, g_ctx.sc);
  if (init) bSwapInit(b);
  else bSwap(b);
}
static void bSelf(lsg_batch *b, const OpParams &op) {
  launchSelfOpSeg(g_ctx.s, b->D, b->cur.b, b->tmp.b, segCtx(b), g_ws->w, op, g_ctx.sc);
  bSwap(b);
}
static void bExpand(lsg_batch *b, bool init, int width) {
  int &w = init ? b->iwidth : b->width;
  if (width <= w) return;
  const double totalLimit = width * 0.5;
  const int startWidth = w;
  for (int cyc = 0; cyc < width - startWidth; ++cyc) {
    OpParams op{};
    op.kind = OP_EXPAND;
    op.a = (startWidth + cyc + 1) * 0.5;
    op.b = totalLimit;
    bStar(b, init, op);
  }
  w = width;
}
static void bRebuild(lsg_batch *b) {
  OpParams op{};
  op.kind = OP_REBUILD;
  op.a = 1.0;
  bStar(b, false, op);
  b->width = 2;
}
static void bStageFE(lsg_batch *b, const double *maxDt, double *dtOut) {
  bExpand(b, false, prepareWidth(b->spatial));
  launchRatesSeg(g_ctx.s, b->D, b->cur.b, segCtx(b), b->rp, b->rates, g_ctx.sc, b->segMax, maxDt, dtOut, g_ws->w);
  OpParams op{};
  op.kind = OP_UPDATE;
  op.rates = b->rates;
  op.dt = dtOut;
  bSelf(b, op);
  b->width = 1;
  bRebuild(b);
}
static void bCombine(lsg_batch *b, double wT, double wS) {
  int steps = (b->temporal == 1) ? 2 : (b->temporal == 2 ? 3 : 1);
  int ew = (int)std::ceil(2.0 * steps * b->cfl + 1);
  bExpand(b, false, ew);
  bExpand(b, true, ew);
  launchBooleanSeg(g_ctx.s, b->D, b->cur.b, b->icur.b, b->tmp.b, segCtx(b), g_ws->w, B_COMBINE, wT, wS, g_ctx.sc);
  bSwap(b);
  OpParams p{};
  p.kind = OP_PRUNE_RSZ;
  bSelf(b, p);
  p.kind = OP_PRUNE;
  bSelf(b, p);
  b->width = 2;
  bRebuild(b);
}
static void bEnqueueStep(lsg_batch *b) {
  const uint32_t S = (uint32_t)b->doms.size();
  launchStepBeginSeg(g_ctx.s, g_ctx.sc, b->segMaxDt, S);
  launchCopyBandWords(g_ctx.s, b->cur.b, b->save.b, b->W);
  double *dt0 = b->segDt, *dt1 = b->segDt + S, *dt2 = b->segDt + 2 * S;
  if (b->temporal == 0) {
    bStageFE(b, b->segMaxDt, dt0);
  } else {
    launchCopyBandWords(g_ctx.s, b->cur.b, b->icur.b, b->W);
    b->iwidth = b->width;
    bStageFE(b, b->segMaxDt, dt0);
    bStageFE(b, dt0, dt1);
    if (b->temporal == 1) {
      bCombine(b, 0.5, 0.5);
    } else {
      bCombine(b, 0.75, 0.25);
      bStageFE(b, dt0, dt2);
      bCombine(b, 1.0 / 3.0, 2.0 / 3.0);
    }
  }
  launchStepTimeSeg(g_ctx.s, g_ctx.sc, b->segDt, S, b->temporal, b->segStepTime, b->segAdvTime);
  launchFrameCheckSeg(g_ctx.s, b->cur.b, segCtx(b), b->D, REFRAME_SLACK, g_ctx.sc);
}

static uint64_t batchKey(lsg_batch *b) {
  uint64_t h = 1469598103934665603ull;
  auto add = [&](const void *p, size_t n) { h = fnv(h, p, n); };
  for (BStore *s : {&b->cur, &b->tmp, &b->icur, &b->itmp, &b->save}) add(&s->b, sizeof(BandBufs));
  add(&b->width, sizeof(int));
  add(&g_ws->w, sizeof(Workspace));
  add(&b->rp, sizeof(RateParams));
  add(&b->temporal, sizeof(int));
  add(&b->segsD, sizeof(void *));
  add(&b->W, sizeof(int64_t));
  add(&b->scanBlocks, sizeof(int));
  return h;
}

static int batchStepOnce(lsg_batch *b);
static int batchStep(lsg_batch *b) {
  const int e = batchStepOnce(b);
  setScanBlocks(defaultScanBlocks());
  return e;
}
static int batchStepOnce(lsg_batch *b) {
  for (int attempt = 0; attempt < 6; ++attempt) {
    setScanBlocks(b->scanBlocks);
    const int widthBefore = b->width;
    bool launched = false;
    if (b->useGraphs) {
      const uint64_t key = batchKey(b);
      lsg_batch::Graph *g = nullptr;
      for (auto &c : b->graphs)
        if (c.exec && c.key == key) g = &c;
      if (!g) {
        g = &b->graphs[b->nextGraph];
        b->nextGraph ^= 1;
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        *g = lsg_batch::Graph{};
        hipGraph_t graph = nullptr;
        bool ok = hipStreamBeginCapture(g_ctx.s, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
          bEnqueueStep(b);
          ok = hipStreamEndCapture(g_ctx.s, &graph) == hipSuccess && graph;
        }
        if (ok) ok = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
        if (!ok) {
          (void)hipGetLastError();
          b->useGraphs = false;
          b->dropGraphs();
          return fail(-31, "HIP graph capture failed; retry the call (direct launches are used from now on)");
        }
        g->key = key;
        g->cur = b->cur;
        g->tmp = b->tmp;
        g->icur = b->icur;
        g->itmp = b->itmp;
        g->width = b->width;
        g->iwidth = b->iwidth;
      } else {
        b->cur = g->cur;
        b->tmp = g->tmp;
        b->icur = g->icur;
        b->itmp = g->itmp;
        b->width = g->width;
        b->iwidth = g->iwidth;
      }
      LSG_HIP(hipGraphLaunch(g->exec, g_ctx.s));
      launched = true;
    }
    if (!launched) bEnqueueStep(b);
    LSG_HIP(hipMemcpyAsync(b->hostSc, g_ctx.sc, sizeof(DevScalars), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    const DevScalars &hs = *b->hostSc;
    if (lookbackFailed(hs)) {  // repeat the step from the saved band
      b->dropGraphs();
      launchCopyBandWords(g_ctx.s, b->save.b, b->cur.b, b->W);
      b->width = widthBefore;
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      continue;
    }
    if (!hs.overflow) {
      b->pointUpdates += (double)(hs.totalActive - hs.stepActive0);
      ++b->steps;
      if (hs.reframe) {
        // move the frames of the level sets that came close to an edge
        b->dropGraphs();
        if (int e = batchWriteBack(b)) return e;
        for (lsg_domain *d : b->doms)
          if (int e = maybeReframe(d, false)) return e;
        return batchAssemble(b);
      }
      return 0;
    }
    // grow and retry from the saved band
    b->dropGraphs();
    const size_t cap = b->cur.b.cap * 2;
    BStore n;
    if (int e = n.alloc(b->W, cap)) return e;
    launchCopyBandWords(g_ctx.s, b->save.b, n.b, b->W);
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    b->cur.release();
    b->cur = n;
    n.b = BandBufs{};
    b->width = widthBefore;
    for (BStore *s : {&b->tmp, &b->save}) if (int e = s->alloc(b->W, cap)) return e;
    if (b->temporal != 0)
      for (BStore *s : {&b->icur, &b->itmp}) if (int e = s->alloc(b->W, cap)) return e;
    if (b->rates) (void)hipFree(b->rates);
    LSG_HIP(hipMalloc(&b->rates, cap * sizeof(double)));
    b->ratesCap = cap;
    if (int e = ensureBatchWs(cap, b->W)) return e;
    // a failed step still accumulated per-level-set times: undo from the saved copy is not
    // needed since stepTime is only added by the completed step's last kernel
  }
  return fail(-30, "band capacity overflow persists after retries");
}

static int batchPrepare(lsg_batch *b) {
  if (b->doms.empty()) return fail(-1, "empty batch");
  for (size_t k = 0; k < b->doms.size(); ++k)
    if (!b->velSet[k]) return fail(-2, "batch: a level set has no velocity");
  const int sid = schemeIdOf(b->spatial);
  if (prepareWidth(b->spatial) < 0 ||
      !(sid == S_EO || sid == S_WENO3 || sid == S_WENO5 || sid == S_LF || sid == S_LLLF || sid == S_LLF))
    return fail(-4, "batch: the Engquist-Osher, Lax-Friedrichs (global, local, local-local) and WENO schemes are "
                    "supported");
  if (!(b->cfl > 0)) return fail(-6, "batch: time step ratio must be positive");
  b->rp = RateParams{};
  b->rp.spatial = b->spatial;
  b->rp.order = schemeOrder(b->spatial);
  b->rp.cfl = b->cfl;
  b->rp.alphaFactor = 1.0;
  b->rp.checkDissipation = 1;
  b->rp.calcNormals = 1;  // WENO: dissipationAlpha (1) as calcNormal (lsAdvect.hpp:670-676)
  if (!b->hostSc) LSG_HIP(hipHostMalloc((void **)&b->hostSc, sizeof(DevScalars), hipHostMallocDefault));
  if (!b->assembled) {
    for (lsg_domain *d : b->doms)
      if (!d->framed) return fail(-7, "batch: level set has no data");
    if (int e = batchAssemble(b)) return e;
  }
  b->rp.segAlpha = b->segAlpha;
  b->rp.segGA = b->segGA;
  return 0;
}

}  // namespace lsg

// ================================================================ C ABI ===

extern "C" {

const char *lsg_last_error(void) { return g_err.c_str(); }

// The grid of a level set (hrleGrid's index range per boundary condition:
// infinite directions reach the grid end points, a periodic direction ends
// one point before its upper bound).
static int setGrid(lsg_domain *d, int D, const int32_t *gmin, const int32_t *gmax, const int *bc, double dx) {
  GridInfo g{};
  for (int k = 0; k < 3; ++k) {
    g.lo[k] = g.hi[k] = 0;
    g.bc[k] = REFLECTIVE;
  }
  for (int k = 0; k < D; ++k) {
    const int b = bc[k];
    g.bc[k] = b;
    switch (b) {
    case INFINITE: g.lo[k] = -INF_IDX; g.hi[k] = INF_IDX; break;
    case PERIODIC: g.lo[k] = gmin[k]; g.hi[k] = gmax[k] - 1; break;
    case POS_INFINITE: g.lo[k] = gmin[k]; g.hi[k] = INF_IDX; break;
    case NEG_INFINITE: g.lo[k] = -INF_IDX; g.hi[k] = gmax[k]; break;
    case REFLECTIVE: g.lo[k] = gmin[k]; g.hi[k] = gmax[k]; break;
    default: return fail(-1, "unknown boundary condition");
    }
  }
  g.D = D;
  g.dx = dx;
  d->D = D;
  d->g = g;
  for (int k = 0; k < 3; ++k) {
    d->umin[k] = (k < D) ? gmin[k] : 0;
    d->umax[k] = (k < D) ? gmax[k] : 0;
  }
  return 0;
}

int lsg_domain_create(int D, const int32_t *gmin, const int32_t *gmax, const int *bc, double grid_delta,
                      lsg_domain **out) {
  if (D != 2 && D != 3) return fail(-1, "dimension must be 2 or 3");
  if (!(grid_delta > 0)) return fail(-1, "gridDelta must be positive");
  if (int e = ctxInit()) return e;
  auto *d = new lsg_domain();
  if (int e = setGrid(d, D, gmin, gmax, bc, grid_delta)) {
    delete d;
    return e;
  }
  // empty level set: one positive undefined run (Domain constructor, lsDomain.hpp:76-77)
  std::vector<int32_t> idx(D);
  for (int k = 0; k < D; ++k) idx[k] = d->g.lo[k];
  std::vector<double> val(1, POS);
  if (int e = uploadRuns(d, idx, val, 1)) {
    delete d;
    return e;
  }
  *out = d;
  return 0;
}

int lsg_domain_destroy(lsg_domain *d) {
  delete d;
  return 0;
}

int lsg_domain_deep_copy(lsg_domain *dst, const lsg_domain *src) {
  if (int e = copyDomain(dst, src)) return e;
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

int lsg_domain_insert_points(lsg_domain *d, int64_t n, const int32_t *idx, const double *val, int width) {
  std::vector<int32_t> pi(idx, idx + n * d->D);
  std::vector<double> pv(val, val + n);
  std::vector<int32_t> ri;
  std::vector<double> rv;
  runsFromPoints(d, pi, pv, ri, rv);
  return uploadRuns(d, ri, rv, width);
}

int lsg_domain_set_runs(lsg_domain *d, int64_t n, const int32_t *idx, const double *val, int width) {
  std::vector<int32_t> ri(idx, idx + n * d->D);
  std::vector<double> rv(val, val + n);
  return uploadRuns(d, ri, rv, width);
}

int lsg_domain_num_runs(lsg_domain *d, int64_t *n) {
  uint32_t h = 0;
  LSG_HIP(hipMemcpyAsync(&h, d->cur.b.n, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  *n = h;
  return 0;
}

// ------------------------------------------------------------ point data ---
static int pdIndex(const lsg_domain *d, int comps, int i) {
  int k = 0;
  for (size_t j = 0; j < d->pd.size(); ++j)
    if (d->pd[j].comps == comps && k++ == i) return (int)j;
  return -1;
}
int lsg_domain_point_data_insert(lsg_domain *d, const char *label, int components, const double *values, int64_t n,
                                 int replace) {
  if (components != 1 && components != 3) return fail(-1, "point data: components must be 1 or 3");
  if (!label || n < 0 || (n > 0 && !values)) return fail(-1, "point data: invalid arguments");
  if (int e = ctxInit()) return e;
  int j = -1;
  if (replace)
    for (size_t q = 0; q < d->pd.size() && j < 0; ++q)
      if (d->pd[q].comps == components && d->pd[q].label == label) j = (int)q;
  if (j < 0) {
    d->pd.emplace_back();
    j = (int)d->pd.size() - 1;
    d->pd[j].label = label;
    d->pd[j].comps = components;
  }
  size_t cap = std::max<size_t>((size_t)n, d->cur.b.cap);
  for (const PdArray &a : d->pd) cap = std::max(cap, a.cap);
  if (int e = pdEnsure(d, std::max<size_t>(cap, 1))) return e;
  PdArray &a = d->pd[j];
  LSG_HIP(hipMemsetAsync(a.buf, 0, a.cap * a.comps * sizeof(double), g_ctx.s));
  if (n > 0)
    LSG_HIP(hipMemcpyAsync(a.buf, values, (size_t)n * components * sizeof(double), hipMemcpyHostToDevice, g_ctx.s));
  uint32_t len = 0;
  LSG_HIP(hipMemcpyAsync(&len, d->pdLen, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  if ((uint32_t)n > len) {
    len = (uint32_t)n;
    LSG_HIP(hipMemcpy(d->pdLen, &len, sizeof(uint32_t), hipMemcpyHostToDevice));
  }
  return 0;
}
int lsg_domain_point_data_count(const lsg_domain *d, int components) {
  int k = 0;
  for (const PdArray &a : d->pd) k += (a.comps == components) ? 1 : 0;
  return k;
}
int lsg_domain_point_data_label(const lsg_domain *d, int components, int i, char *buf, int len) {
  const int j = pdIndex(d, components, i);
  if (j < 0) return fail(-1, "point data: no such array");
  if (buf && len > 0) snprintf(buf, (size_t)len, "%s", d->pd[j].label.c_str());
  return (int)d->pd[j].label.size();
}
// Tuples per array: every array holds one value (or vector) per point id
// of the last translation (arrays inserted since hold what was given,
// padded with zeros up to the longest).
int64_t lsg_domain_point_data_size(const lsg_domain *d, int components, int i) {
  if (pdIndex(d, components, i) < 0) return fail(-1, "point data: no such array");
  if (int e = ctxInit()) return e;
  uint32_t len = 0;
  LSG_HIP(hipMemcpyAsync(&len, d->pdLen, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return (int64_t)len;
}
int lsg_domain_point_data_get(const lsg_domain *d, int components, int i, double *out, int64_t n) {
  const int j = pdIndex(d, components, i);
  if (j < 0) return fail(-1, "point data: no such array");
  const int64_t len = lsg_domain_point_data_size(d, components, i);
  if (len < 0) return (int)len;
  if (n < len) return fail(-1, "point data: output buffer too small");
  if (len > 0)
    LSG_HIP(hipMemcpyAsync(out, d->pd[j].buf, (size_t)len * components * sizeof(double), hipMemcpyDeviceToHost,
                           g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}
int lsg_domain_point_data_clear(lsg_domain *d) {
  pdClear(d);
  return 0;
}

// ------------------------------------------------------- serialisation ---
// Domain::serialize / deserialize (lsDomain.hpp:188-268).  The envelope is
// the reference's: "lsDomain", the format version byte (0), the grid, the
// HRLE domain, the level-set width as uint32 and a byte saying whether
// point data follows.  ViennaHRLE's grid and domain serialisers and
// ViennaCore's PointData serialiser write the three inner sections; they
// are not part of the reference, so these sections have a layout of this
// library (native byte order):
//   grid:  "lsgGrid" NUL, uint8 D, per direction int32 minimum, int32
//          maximum (the bounds as given), uint8 boundary condition, then
//          float64 gridDelta
//   runs:  "lsgRuns" NUL, uint64 n, n x D int32 run starts (HRLE order,
//          entry 0 at the grid minimum), n float64 run values
//   point data: uint32 number of scalar arrays, each uint32 label length,
//          the label, uint64 tuples, the float64 values; then the vector
//          arrays alike with 3 float64 per tuple
extern "C++" {
namespace {
struct ByteOut {
  std::vector<uint8_t> b;
  void put(const void *p, size_t n) {
    const uint8_t *c = (const uint8_t *)p;
    b.insert(b.end(), c, c + n);
  }
  template <class T> void val(T x) { put(&x, sizeof(T)); }
};
struct ByteIn {
  const uint8_t *p;
  size_t n, o;
  bool get(void *dst, size_t k) {
    if (k > n - o) return false;
    std::memcpy(dst, p + o, k);
    o += k;
    return true;
  }
  template <class T> bool val(T &x) { return get(&x, sizeof(T)); }
};
}  // namespace
static constexpr char DOMAIN_FORMAT_VERSION = 0;  // LS_DOMAIN_SERIALIZATION_VERSION (lsDomain.hpp:16)

static int serializeDomain(lsg_domain *d, std::vector<uint8_t> &bytes) {
  ByteOut o;
  o.put("lsDomain", 8);
  o.val<char>(DOMAIN_FORMAT_VERSION);
  o.put("lsgGrid", 8);
  o.val<uint8_t>((uint8_t)d->D);
  for (int k = 0; k < d->D; ++k) {
    o.val<int32_t>(d->umin[k]);
    o.val<int32_t>(d->umax[k]);
    o.val<uint8_t>((uint8_t)d->g.bc[k]);
  }
  o.val<double>(d->g.dx);
  std::vector<int32_t> idx;
  std::vector<double> val;
  if (int e = downloadRuns(d, idx, val)) return e;
  o.put("lsgRuns", 8);
  o.val<uint64_t>((uint64_t)val.size());
  o.put(idx.data(), idx.size() * sizeof(int32_t));
  o.put(val.data(), val.size() * sizeof(double));
  o.val<uint32_t>((uint32_t)d->width);
  const char hasPointData = d->pd.empty() ? 0 : 1;
  o.val<char>(hasPointData);
  if (hasPointData) {
    for (int comps : {1, 3}) {
      const int na = lsg_domain_point_data_count(d, comps);
      o.val<uint32_t>((uint32_t)na);
      for (int i = 0; i < na; ++i) {
        const std::string &label = d->pd[pdIndex(d, comps, i)].label;
        o.val<uint32_t>((uint32_t)label.size());
        o.put(label.data(), label.size());
        const int64_t len = lsg_domain_point_data_size(d, comps, i);
        if (len < 0) return (int)len;
        std::vector<double> v((size_t)len * comps);
        if (int e = lsg_domain_point_data_get(d, comps, i, v.data(), len)) return e;
        o.val<uint64_t>((uint64_t)len);
        o.put(v.data(), v.size() * sizeof(double));
      }
    }
  }
  bytes.swap(o.b);
  return 0;
}
}  // extern "C++"

int lsg_domain_serialize(lsg_domain *d, uint8_t *buf, int64_t cap, int64_t *size) {
  if (int e = ctxInit()) return e;
  std::vector<uint8_t> bytes;
  if (int e = serializeDomain(d, bytes)) return e;
  if (size) *size = (int64_t)bytes.size();
  if (!buf) return 0;
  if (cap < (int64_t)bytes.size()) return fail(-1, "serialize: buffer too small");
  std::memcpy(buf, bytes.data(), bytes.size());
  return 0;
}

// The stream is read completely before the level set changes: a stream
// that is cut short or malformed leaves it as it was.  As in the
// reference, point data the level set holds stays when the stream has
// none (lsDomain.hpp:262-266).
int lsg_domain_deserialize(lsg_domain *d, const uint8_t *buf, int64_t size) {
  if (int e = ctxInit()) return e;
  if (!buf || size < 0) return fail(-1, "deserialize: no data");
  ByteIn in{buf, (size_t)size, 0};
  char tag[8];
  if (!in.get(tag, 8) || std::memcmp(tag, "lsDomain", 8) != 0)
    return fail(-60, "Reading Domain from stream failed. Header could not be found.");
  char version = 0;
  if (!in.val(version)) return fail(-61, "Reading Domain from stream failed: stream ends early.");
  if (version > DOMAIN_FORMAT_VERSION)
    return fail(-62, "Reading Domain of version " + std::to_string((int)version) + " with reader of version " +
                         std::to_string((int)DOMAIN_FORMAT_VERSION) + " failed.");
  const auto early = [] { return fail(-61, "Reading Domain from stream failed: stream ends early or is malformed."); };
  uint8_t D = 0;
  if (!in.get(tag, 8) || std::memcmp(tag, "lsgGrid", 8) != 0 || !in.val(D)) return early();
  if ((int)D != d->D) return fail(-63, "Reading Domain from stream failed: the stream holds a " + std::to_string(D) +
                                           "D level set, this one is " + std::to_string(d->D) + "D.");
  int32_t gmin[3] = {0, 0, 0}, gmax[3] = {0, 0, 0};
  int bc[3] = {0, 0, 0};
  for (int k = 0; k < D; ++k) {
    uint8_t b = 0;
    if (!in.val(gmin[k]) || !in.val(gmax[k]) || !in.val(b)) return early();
    bc[k] = b;
  }
  double dx = 0;
  if (!in.val(dx) || !(dx > 0)) return early();
  uint64_t n = 0;
  if (!in.get(tag, 8) || std::memcmp(tag, "lsgRuns", 8) != 0 || !in.val(n)) return early();
  if (n == 0 || n > (uint64_t)(size / (int64_t)(D * sizeof(int32_t) + sizeof(double)))) return early();
  std::vector<int32_t> idx((size_t)n * D);
  std::vector<double> val((size_t)n);
  uint32_t width = 0;
  char hasPointData = 0;
  if (!in.get(idx.data(), idx.size() * sizeof(int32_t)) || !in.get(val.data(), val.size() * sizeof(double)) ||
      !in.val(width) || !in.val(hasPointData))
    return early();
  struct Arr {
    std::string label;
    int comps;
    std::vector<double> v;
  };
  std::vector<Arr> arrays;
  if (hasPointData == 1) {
    for (int comps : {1, 3}) {
      uint32_t na = 0;
      if (!in.val(na)) return early();
      for (uint32_t i = 0; i < na; ++i) {
        uint32_t len = 0;
        if (!in.val(len) || len > size) return early();
        std::string label(len, '\0');
        uint64_t tuples = 0;
        if (!in.get(&label[0], len) || !in.val(tuples) || tuples > (uint64_t)size / (8u * comps)) return early();
        std::vector<double> v((size_t)tuples * comps);
        if (!in.get(v.data(), v.size() * sizeof(double))) return early();
        arrays.push_back(Arr{label, comps, std::move(v)});
      }
    }
  }
  // build the level set on its own, then take it over
  std::unique_ptr<lsg_domain> t(new lsg_domain());
  if (int e = setGrid(t.get(), D, gmin, gmax, bc, dx)) return e;
  if (int e = uploadRuns(t.get(), idx, val, (int)width)) return e;
  for (const Arr &a : arrays)
    if (int e = lsg_domain_point_data_insert(t.get(), a.label.c_str(), a.comps, a.v.data(), (int64_t)(a.v.size() / a.comps), 0))
      return e;
  d->D = t->D;
  d->g = t->g;
  for (int k = 0; k < 3; ++k) {
    d->umin[k] = t->umin[k];
    d->umax[k] = t->umax[k];
  }
  d->f = t->f;
  d->framed = t->framed;
  d->width = t->width;
  if (int e = ensureCur(d, t->cur.nwAlloc, t->cur.b.cap)) return e;
  launchCopyBand(g_ctx.s, t->cur.b, d->cur.b, t->f);
  d->hostN = t->hostN;
  if (hasPointData == 1) {
    pdClear(d);
    if (int e = pdCopy(d, t.get())) return e;
  }
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

int lsg_domain_get_grid(const lsg_domain *d, int *D, int32_t *gmin, int32_t *gmax, int *bc, double *dx) {
  if (D) *D = d->D;
  for (int k = 0; k < d->D; ++k) {
    if (gmin) gmin[k] = d->umin[k];
    if (gmax) gmax[k] = d->umax[k];
    if (bc) bc[k] = d->g.bc[k];
  }
  if (dx) *dx = d->g.dx;
  return 0;
}
// ------------------------------------------------------ surface & Chamfer ---
struct SurfBuf {
  double *pts = nullptr;
  uint32_t *start = nullptr, *total = nullptr;
  uint32_t n = 0;  // surface points
  int release() {
    for (void *p : {(void *)pts, (void *)start, (void *)total}) if (p) (void)hipFree(p);
    return 0;
  }
};
static int surfaceOf(lsg_domain *d, SurfBuf &s) {
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  uint32_t ne = 0;
  LSG_HIP(hipMemcpyAsync(&ne, d->cur.b.n, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  const size_t maxPts = (size_t)ne * d->D + 1;
  LSG_HIP(hipMalloc(&s.pts, maxPts * d->D * sizeof(double)));
  LSG_HIP(hipMalloc(&s.start, ((size_t)ne + 1) * sizeof(uint32_t)));
  LSG_HIP(hipMalloc(&s.total, sizeof(uint32_t)));
  launchSurfacePoints(g_ctx.s, d->D, d->cur.b, d->f, d->g, g_ws->w, s.pts, s.start, s.total);
  LSG_HIP(hipMemcpyAsync(&s.n, s.total, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}
int lsg_surface_points(lsg_domain *d, double *out, int64_t cap, int64_t *n) {
  if (int e = ctxInit()) return e;
  if (!d->framed) return fail(-7, "surface points: level set has no data");
  SurfBuf s;
  if (int e = surfaceOf(d, s)) return s.release(), e;
  *n = s.n;
  int rc = 0;
  if (out) {
    if ((int64_t)s.n > cap) rc = fail(-1, "surface points: output buffer too small");
    else if (s.n) rc = hipMemcpy(out, s.pts, (size_t)s.n * d->D * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess ? 0 : fail(-2, "copy failed");
  }
  s.release();
  return rc;
}
// Nearest distances of the surface points of `from` to the surface of `to`.
static int nearestAll(lsg_domain *from, const SurfBuf &sf, lsg_domain *to, const SurfBuf &st, std::vector<double> &dist) {
  dist.assign(sf.n, 0.0);
  if (sf.n == 0) return 0;
  int maxShell = 2;
  for (int k = 0; k < to->D; ++k) {
    const int64_t lo = std::min<int64_t>(from->f.o[k], to->f.o[k]);
    const int64_t hi = std::max<int64_t>((int64_t)from->f.o[k] + from->f.ext[k], (int64_t)to->f.o[k] + to->f.ext[k]);
    maxShell = std::max<int>(maxShell, (int)(hi - lo) + 2);
  }
  double *dd = nullptr;
  LSG_HIP(hipMalloc(&dd, (size_t)sf.n * sizeof(double)));
  launchNearest(g_ctx.s, to->D, sf.pts, sf.total, sf.n, to->cur.b, to->f, to->g, st.start, st.pts, maxShell, dd);
  hipError_t e = hipMemcpyAsync(dist.data(), dd, (size_t)sf.n * sizeof(double), hipMemcpyDeviceToHost, g_ctx.s);
  if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.s);
  (void)hipFree(dd);
  if (e != hipSuccess) return fail(-2, hipGetErrorString(e));
  return 0;
}
static int afterOp(lsg_domain *d);

// The level set a judge reads: a copy expanded to minWidth when the level
// set is narrower, else the level set itself (CompareChamfer: width 2, a zero
// crossing needs both edge end points defined, lsCompareChamfer.hpp:118-134;
// CompareVolume: width 3, lsCompareVolume.hpp:195-220).
static int judgeInput(lsg_domain *d, int minWidth, std::unique_ptr<lsg_domain> &copy, lsg_domain *&use) {
  use = d;
  if (d->width >= minWidth) return 0;
  copy.reset(new lsg_domain());
  if (int e = copyDomain(copy.get(), d)) return e;
  if (int e = ensureStores(copy.get(), copy->cur.b.cap)) return e;
  doExpand(copy.get(), minWidth);
  if (int e = afterOp(copy.get())) return e;
  use = copy.get();
  return 0;
}

// CompareChamfer::apply (lsCompareChamfer.hpp:104-245); out = forward,
// backward, Chamfer, RMS and max distance; counts = target, sample points.
// Sums run on the host in point order, as the reference's loops do.
int lsg_compare_chamfer(lsg_domain *target0, lsg_domain *sample0, double *out, int64_t *counts) {
  if (int e = ctxInit()) return e;
  if (target0->D != sample0->D) return fail(-1, "CompareChamfer: dimensions differ");
  if (!target0->framed || !sample0->framed) return fail(-7, "CompareChamfer: level set has no data");
  if (target0->g.dx != sample0->g.dx) return fail(-1, "CompareChamfer: grid deltas differ");
  std::unique_ptr<lsg_domain> tcopy, scopy;
  lsg_domain *target = nullptr, *sample = nullptr;
  if (int e = judgeInput(target0, 2, tcopy, target)) return e;
  if (int e = judgeInput(sample0, 2, scopy, sample)) return e;
  SurfBuf st, ss;
  int rc = surfaceOf(target, st);
  if (!rc) rc = surfaceOf(sample, ss);
  std::vector<double> fwd, bwd;
  if (!rc && st.n && ss.n) rc = nearestAll(target, st, sample, ss, fwd);
  if (!rc && st.n && ss.n) rc = nearestAll(sample, ss, target, st, bwd);
  counts[0] = st.n;
  counts[1] = ss.n;
  st.release();
  ss.release();
  if (rc) return rc;
  if (fwd.empty() || bwd.empty()) {
    for (int k = 0; k < 5; ++k) out[k] = std::numeric_limits<double>::infinity();
    return 0;
  }
  double sf = 0, sf2 = 0, mf = 0, sb = 0, sb2 = 0, mb = 0;
  for (double x : fwd) {
    sf += x;
    sf2 += x * x;
    mf = std::max(mf, x);
  }
  for (double x : bwd) {
    sb += x;
    sb2 += x * x;
    mb = std::max(mb, x);
  }
  out[0] = sf / (double)fwd.size();
  out[1] = sb / (double)bwd.size();
  out[2] = (sf + sb) / (double)(fwd.size() + bwd.size());
  out[3] = std::sqrt((sf2 + sb2) / (double)(fwd.size() + bwd.size()));
  out[4] = std::max(mf, mb);
  return 0;
}

// CompareVolume::apply (lsCompareVolume.hpp:176-289) without custom
// increments or mesh output.  The box spans the grid along bounded axes
// (reflective: cells lo..hi-1; periodic: lo..hi, the last cell wraps) and,
// along the others, the cells touching a run start of either level set (the
// grid-minimum entry excluded), which restates ViennaHRLE's run breaks.
int lsg_compare_volume(lsg_domain *target0, lsg_domain *sample0, uint64_t *cells, double *volume) {
  if (int e = ctxInit()) return e;
  if (target0->D != sample0->D) return fail(-1, "CompareVolume: dimensions differ");
  if (!target0->framed || !sample0->framed) return fail(-7, "CompareVolume: level set has no data");
  if (target0->g.dx != sample0->g.dx)
    return fail(-1, "CompareVolume: grid delta mismatch; the grid deltas of the two level sets must be equal");
  for (int d = 0; d < target0->D; ++d)
    if (target0->g.lo[d] != sample0->g.lo[d] || target0->g.hi[d] != sample0->g.hi[d] ||
        target0->g.bc[d] != sample0->g.bc[d])
      return fail(-1, "CompareVolume: the level sets must share their grid bounds and boundary conditions");
  std::unique_ptr<lsg_domain> tcopy, scopy;
  lsg_domain *A = nullptr, *B = nullptr;
  if (int e = judgeInput(target0, 3, tcopy, A)) return e;
  if (int e = judgeInput(sample0, 3, scopy, B)) return e;
  const int D = A->D;
  int64_t bmin[3], bmax[3];
  bool any = false;
  for (lsg_domain *d : {A, B}) {
    launchBandStats(g_ctx.s, d->cur.b, d->f, g_ctx.stat, (int32_t *)(g_ctx.stat + 1), 1u);
    int32_t h[7];
    LSG_HIP(hipMemcpyAsync(h, g_ctx.stat, 7 * sizeof(int32_t), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    if (h[1] > h[4]) continue;  // no in-frame entry after entry 0
    for (int k = 0; k < D; ++k) {
      const int64_t a = (int64_t)h[1 + k] + d->f.o[k], b = (int64_t)h[4 + k] + d->f.o[k];
      bmin[k] = any ? std::min(bmin[k], a) : a;
      bmax[k] = any ? std::max(bmax[k], b) : b;
    }
    any = true;
  }
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  bool empty = false;
  for (int k = 0; k < D; ++k) {
    if (A->g.bc[k] == PERIODIC) {
      lo[k] = A->g.lo[k];
      hi[k] = A->g.hi[k];
    } else if (A->g.bc[k] == REFLECTIVE) {
      lo[k] = A->g.lo[k];
      hi[k] = (int64_t)A->g.hi[k] - 1;
    } else if (any) {
      lo[k] = bmin[k] - 1;
      hi[k] = bmax[k];
    } else {
      empty = true;
    }
    if (hi[k] < lo[k]) empty = true;
  }
  unsigned long long n = 0;
  if (!empty) {
    unsigned long long *dn = nullptr;
    LSG_HIP(hipMalloc(&dn, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(dn, 0, sizeof(unsigned long long), g_ctx.s);
    if (e == hipSuccess) {
      launchCompareVolume(g_ctx.s, D, A->cur.b, A->f, B->cur.b, B->f, A->g, lo, hi, dn);
      e = hipMemcpyAsync(&n, dn, sizeof(n), hipMemcpyDeviceToHost, g_ctx.s);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_ctx.s);
    (void)hipFree(dn);
    if (e != hipSuccess) return fail(-100, std::string("HIP error: ") + hipGetErrorString(e));
  }
  if (cells) *cells = n;
  if (volume) *volume = (double)n * std::pow(A->g.dx, D);
  return 0;
}

int lsg_mark_void_points(lsg_domain *d) {
  if (int e = ctxInit()) return e;
  if (!d->framed) return fail(-7, "MarkVoidPoints: level set has no data");
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  RateParams rp{};
  if (int e = markVoids(d, rp)) return e;
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

int lsg_domain_get_runs(lsg_domain *d, int32_t *idx, double *val, int64_t cap) {
  std::vector<int32_t> ri;
  std::vector<double> rv;
  if (int e = downloadRuns(d, ri, rv)) return e;
  if ((int64_t)rv.size() > cap) return fail(-1, "buffer too small");
  std::memcpy(idx, ri.data(), ri.size() * sizeof(int32_t));
  std::memcpy(val, rv.data(), rv.size() * sizeof(double));
  return 0;
}

int lsg_domain_num_points(lsg_domain *d, int64_t *n) {
  uint32_t nd;
  int32_t a[3], b[3];
  bool any;
  if (int e = bandStats(d, nd, a, b, any)) return e;
  *n = nd;
  return 0;
}

int lsg_domain_width(lsg_domain *d, int *w) {
  *w = d->width;
  return 0;
}
int lsg_domain_set_width(lsg_domain *d, int w) {
  if (w < 1) return fail(-1, "level set width must be at least 1");
  d->width = w;
  return 0;
}

int lsg_make_sphere(lsg_domain *d, const double *origin, double radius) {
  return lsg_make_sphere_ignore_bc(d, origin, radius, 0);
}
int lsg_make_sphere_ignore_bc(lsg_domain *d, const double *origin, double radius, int ignore_mask) {
  std::vector<int32_t> pi;
  std::vector<double> pv;
  sphereBand(d->D, d->g, origin, radius, pi, pv, ignore_mask);
  std::vector<int32_t> ri;
  std::vector<double> rv;
  runsFromPoints(d, pi, pv, ri, rv);
  return uploadRuns(d, ri, rv, 2);
}

int lsg_make_plane(lsg_domain *d, const double *origin, const double *normal) {
  std::vector<int32_t> pi;
  std::vector<double> pv;
  int e = planeBand(d->D, d->g, origin, normal, pi, pv);
  if (e == -1) return fail(-1, "MakeGeometry: plane normal is zero");
  if (e == -2) return fail(-1, "Planes can only be created with one Infinite Boundary Condition. More than one found!");
  if (e == -3) return fail(-1, "MakeGeometry: Plane cannot be parallel to Infinite Boundary direction!");
  if (e == -4) return fail(-1, "Planes require exactly one Infinite Boundary Condition. None found!");
  std::vector<int32_t> ri;
  std::vector<double> rv;
  runsFromPoints(d, pi, pv, ri, rv);
  return uploadRuns(d, ri, rv, 2);
}

int lsg_make_box(lsg_domain *d, const double *mn, const double *mx) { return lsg_make_box_ignore_bc(d, mn, mx, 0); }
int lsg_make_box_ignore_bc(lsg_domain *d, const double *mn, const double *mx, int ignore_mask) {
  std::vector<int32_t> pi;
  std::vector<double> pv;
  boxBand(d->D, d->g, mn, mx, pi, pv, ignore_mask);
  std::vector<int32_t> ri;
  std::vector<double> rv;
  runsFromPoints(d, pi, pv, ri, rv);
  return uploadRuns(d, ri, rv, 2);
}

static int afterOp(lsg_domain *d) {
  uint32_t lb = 0;
  LSG_HIP(hipMemcpyAsync(&lb, &g_ctx.sc->lookbackFail, sizeof(uint32_t), hipMemcpyDeviceToHost, g_ctx.s));
  if (readOverflow(true)) return fail(-40, "band capacity overflow");
  if (lb) {
    DevScalars h{};
    h.lookbackFail = lb;
    (void)lookbackFailed(h);
    return fail(-47, "a single-pass band operation did not complete; the level set is invalid (the multi-kernel "
                     "band operations are used from now on)");
  }
  return maybeReframe(d, false);
}

int lsg_expand(lsg_domain *d, int width) {
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  doExpand(d, width);
  return afterOp(d);
}

// Expand without allocating: the cycles of lsg_expand (doExpand) as
// single-pass star kernels whose flags, statistics and tile status words
// live in the file-scope `scratch` array, the level set's own two band
// stores as input and output.  One synchronisation per cycle reads the
// flags: after an overflow or an unfinished look-back the level set keeps
// the last complete cycle (the width says which).
int lsg_expand_noalloc(lsg_domain *d, int width) {
  if (!g_ctx.init || !d->framed) return fail(-7, "lsg_expand_noalloc: level set has no data");
  int cur = -1;
  LSG_HIP(hipGetDevice(&cur));
  if (cur != g_ctx.dev) return fail(-50, "lsg_expand_noalloc: the library runs on device " + std::to_string(g_ctx.dev));
  if (width <= d->width) return 0;
  if (!d->pd.empty()) return fail(-48, "lsg_expand_noalloc: point data is translated with allocated buffers; use lsg_expand");
  const int64_t nw = d->f.nwords;
  if (nw > NOALLOC_MAX_WORDS)
    return fail(-49, "lsg_expand_noalloc: frame of " + std::to_string(nw) + " bitmap words exceeds the limit of " +
                         std::to_string(NOALLOC_MAX_WORDS));
  if (!d->tmp.b.val || d->tmp.nwAlloc < nw || !d->cur.b.val)
    return fail(-7, "lsg_expand_noalloc: level set has no output store");
  DevScalars *sc = (DevScalars *)g_ctx.scratch;
  uint32_t *stat = (uint32_t *)(g_ctx.scratch + NOALLOC_SC_BYTES);
  uint64_t *tiles = (uint64_t *)(g_ctx.scratch + NOALLOC_SC_BYTES + NOALLOC_STAT_BYTES);
  const int cycles = width - d->width;
  // the expanded band keeps REFRAME_SLACK cells to every frame edge that
  // is not a grid end (the state lsg_expand leaves without moving the
  // frame); each cycle grows the band by at most one cell per axis
  launchBandStats(g_ctx.s, d->cur.b, d->f, stat, (int32_t *)(stat + 1));
  uint32_t h[7];
  LSG_HIP(hipMemcpyAsync(h, stat, sizeof(h), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  const int32_t *bb = (const int32_t *)(h + 1);
  if (bb[0] <= bb[3])
    for (int k = 0; k < d->D; ++k) {
      if (d->g.bounded(k)) continue;
      const int64_t top = (int64_t)d->f.o[k] + d->f.ext[k] - 1;
      const int64_t lo = (d->f.o[k] == d->g.lo[k]) ? INT64_MAX : bb[k];
      const int64_t hi = (top == d->g.hi[k]) ? INT64_MAX : (int64_t)d->f.ext[k] - 1 - bb[3 + k];
      if (std::min(lo, hi) < (int64_t)cycles + REFRAME_SLACK)
        return fail(-44, "lsg_expand_noalloc: the expanded band would come within " + std::to_string(REFRAME_SLACK) +
                             " cells of its frame edge; lsg_expand moves the frame");
    }
  launchClearFlags(g_ctx.s, sc);
  const double totalLimit = width * 0.5;
  const int startWidth = d->width;
  for (int cyc = 0; cyc < cycles; ++cyc) {
    OpParams op{};
    op.kind = OP_EXPAND;
    op.a = (startWidth + cyc + 1) * 0.5;
    op.b = totalLimit;
    launchStarOpScratch(g_ctx.s, d->D, d->cur.b, d->tmp.b, d->f, d->g, op, tiles, sc);
    uint32_t flags[2];
    static_assert(offsetof(DevScalars, reframe) == offsetof(DevScalars, overflow) + 4, "flag layout");
    LSG_HIP(hipMemcpyAsync(flags, &sc->overflow, sizeof(flags), hipMemcpyDeviceToHost, g_ctx.s));
    uint32_t lb = 0;
    LSG_HIP(hipMemcpyAsync(&lb, &sc->lookbackFail, sizeof(lb), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    if (flags[0] || lb) {
      // status words of tiles that published after the launch counter moved
      // on would read as ready in the next launch: clear them all
      if (lb) launchFill32(g_ctx.s, (uint32_t *)tiles, 0u, 2 * NOALLOC_TILES);
      launchClearFlags(g_ctx.s, sc);
      LSG_HIP(hipStreamSynchronize(g_ctx.s));
      d->width = startWidth + cyc;
      if (lb) return fail(-47, "lsg_expand_noalloc: a single-pass band operation did not complete");
      return fail(-40, "lsg_expand_noalloc: band capacity overflow");
    }
    swapCur(d);
  }
  d->width = width;
  return 0;
}
int lsg_reduce(lsg_domain *d, int width) {
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  doReduce(d, width);
  return afterOp(d);
}
int lsg_prune(lsg_domain *d, int rsz) {
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  doPrune(d, rsz != 0);
  return afterOp(d);
}
int lsg_rebuild(lsg_domain *d) {
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  doRebuild(d);
  return afterOp(d);
}
int lsg_boolean(lsg_domain *a, const lsg_domain *b, int op) {
  if (op == 3) {  // INVERT: b is not used (lsBooleanOperation.hpp:214-247)
    if (int e = ctxInit()) return e;
    if (!a->framed) return fail(-7, "BooleanOperation: level set has no data");
    launchNegate(g_ctx.s, a->cur.b);
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    return 0;
  }
  if (op < 0 || op > 2) return fail(-1, "unsupported boolean operation");
  if (int e = commonFrame(a, const_cast<lsg_domain *>(b))) return e;
  if (int e = ensureStores(a, std::max(a->cur.b.cap, b->cur.b.cap))) return e;
  doBoolean(a, b, op, 0, 0);
  return afterOp(a);
}

int lsg_advect_create(lsg_advect **out) {
  *out = new lsg_advect();
  return 0;
}
int lsg_advect_destroy(lsg_advect *a) {
  delete a;
  return 0;
}
int lsg_advect_insert_next_level_set(lsg_advect *a, lsg_domain *d) {
  a->levelSets.push_back(d);
  return 0;
}
int lsg_advect_clear_level_sets(lsg_advect *a) {
  a->levelSets.clear();
  return 0;
}
int lsg_advect_set_dissipation_alpha_builtin(lsg_advect *a, int material, const double *alpha3) {
  if (material < -1 || material >= MAX_LEVEL_SETS) return fail(-1, "material index out of range (-1..7)");
  double *dst = (material < 0) ? a->rp.velAlpha : a->rp.matAlpha[material];
  for (int i = 0; i < 3; ++i) dst[i] = alpha3 ? alpha3[i] : 0.0;
  a->dropGraphs();
  return 0;
}
int lsg_advect_set_velocity_builtin_material(lsg_advect *a, int material, int kind, int dir, int k, double b,
                                             double aPos, double aNeg, const double *vec3) {
  if (material < 0 || material >= MAX_LEVEL_SETS) return fail(-1, "material index out of range (0..7)");
  if (kind < 0 || kind > 2) return fail(-1, "unknown velocity kind");
  if (dir < 0 || dir > 2) return fail(-1, "velocity direction must be 0, 1 or 2");
  RateParams &rp = a->rp;
  rp.matSet[material] = 1;
  rp.matKind[material] = kind;
  rp.matDir[material] = dir;
  rp.matK[material] = k;
  rp.matB[material] = b;
  rp.matAPos[material] = aPos;
  rp.matANeg[material] = aNeg;
  for (int i = 0; i < 3; ++i) rp.matVec[material][i] = vec3 ? vec3[i] : 0.0;
  for (int i = 0; i < 3; ++i) rp.matAlpha[material][i] = 0.0;
  a->dropGraphs();
  return 0;
}
int lsg_advect_set_velocity_host(lsg_advect *a, lsg_scalar_velocity_fn scalar, lsg_vector_velocity_fn vector,
                                 void *user) {
  if (!scalar && !vector) return fail(-1, "lsg_advect_set_velocity_host: no callback");
  a->hostScalar = scalar;
  a->hostVector = vector;
  a->hostUser = user;
  a->dropGraphs();
  a->velSet = true;
  return 0;
}
int lsg_advect_set_dissipation_alpha_host(lsg_advect *a, lsg_dissipation_alpha_fn fn, void *user) {
  a->hostAlpha = fn;
  a->hostAlphaUser = user;
  a->dropGraphs();
  return 0;
}
int lsg_advect_set_velocity_builtin(lsg_advect *a, int kind, int dir, int k, double b, double aPos, double aNeg,
                                    const double *vec3) {
  a->hostScalar = nullptr;
  a->hostVector = nullptr;
  a->hostUser = nullptr;
  a->hostAlpha = nullptr;
  a->hostAlphaUser = nullptr;
  if (kind < 0 || kind > 2) return fail(-1, "unknown velocity kind");
  if (dir < 0 || dir > 2) return fail(-1, "velocity direction must be 0, 1 or 2");
  a->rp.velKind = kind;
  a->rp.velDir = dir;
  a->rp.velK = k;
  a->rp.velB = b;
  a->rp.velAPos = aPos;
  a->rp.velANeg = aNeg;
  for (int i = 0; i < 3; ++i) a->rp.vec[i] = vec3 ? vec3[i] : 0.0;
  for (int m = 0; m < MAX_LEVEL_SETS; ++m) a->rp.matSet[m] = 0;  // a new field replaces per-material functors
  for (int i = 0; i < 3; ++i) a->rp.velAlpha[i] = 0.0;
  a->velSet = true;
  return 0;
}
int lsg_advect_set_spatial_scheme(lsg_advect *a, int s) {
  if (s < 0 || s > 12) return fail(-1, "Advect: Discretization scheme not found.");
  a->spatial = s;
  return 0;
}
int lsg_advect_set_learned_weights(lsg_advect *a, const float *params, int n) {
  if (n != lsg::MLP_NPARAM) return fail(-1, "learned weno5: expected 1347 network parameters");
  if (int e = lsg::ctxInit()) return e;
  std::vector<uint16_t> b(n);
  for (int i = 0; i < n; ++i) b[i] = lsg::bf16Bits(params[i]);
  if (!a->mlpDev) LSG_HIP(hipMalloc(&a->mlpDev, n * sizeof(uint16_t)));
  LSG_HIP(hipMemcpy(a->mlpDev, b.data(), n * sizeof(uint16_t), hipMemcpyHostToDevice));
  return 0;
}
int lsg_advect_set_temporal_scheme(lsg_advect *a, int s) {
  a->temporal = (s == 1 || s == 2) ? s : 0;
  return 0;
}
int lsg_advect_set_advection_time(lsg_advect *a, double t) {
  a->advectionTime = t;
  return 0;
}
int lsg_advect_set_time_step_ratio(lsg_advect *a, double c) {
  a->cfl = c;
  return 0;
}
int lsg_advect_set_single_step(lsg_advect *a, int s) {
  a->singleStep = s;
  return 0;
}
int lsg_advect_set_dissipation_alpha(lsg_advect *a, double v) {
  a->dissipationAlpha = v;
  return 0;
}
int lsg_advect_set_check_dissipation(lsg_advect *a, int c) {
  a->checkDissipation = c;
  return 0;
}
int lsg_advect_set_calculate_normal_vectors(lsg_advect *a, int c) {
  a->calcNormals = c;
  return 0;
}
int lsg_advect_set_ignore_voids(lsg_advect *a, int iv) {
  a->ignoreVoids = iv;
  return 0;
}
int lsg_advect_set_adaptive_time_stepping(lsg_advect *a, int on, unsigned subdivisions) {
  if (subdivisions < 1) {
    fprintf(stderr, "[viennals_gpu] Advect: Adaptive time stepping subdivisions must be at least 1. Setting to 1.\n");
    subdivisions = 1;
  }
  a->rp.adaptive = on ? 1 : 0;
  a->rp.adaptiveFactor = 1.0 / subdivisions;
  a->dropGraphs();
  return 0;
}
int lsg_advect_set_update_point_data(lsg_advect *a, int update) {
  a->updatePointData = update != 0;
  return 0;
}
int lsg_advect_set_save_advection_velocities(lsg_advect *a, int save) {
  a->saveVelocities = save != 0;
  a->dropGraphs();
  return 0;
}
int lsg_advect_prepare_ls(lsg_advect *a) {
  if (a->levelSets.empty()) return fail(-1, "No level sets passed to Advect.");
  int w = prepareWidth(a->spatial);
  if (w < 0) return fail(-1, "Advect: Discretization scheme not found.");
  lsg_domain *d = a->levelSets.back();
  if (int e = ensureStores(d, d->cur.b.cap)) return e;
  doExpand(d, w);
  return afterOp(d);
}
int lsg_advect_apply(lsg_advect *a) {
  if (g_inVelocityUpdate) return fail(-1, "Advect::apply called from a velocity update callback");
  return advectApply(a);
}
int lsg_advect_set_velocity_update_callback(lsg_advect *a, lsg_velocity_update_fn fn, void *user) {
  a->velUpdate = fn;
  a->velUpdateUser = fn ? user : nullptr;
  return 0;
}
int lsg_advect_get_advected_time(lsg_advect *a, double *t) {
  *t = a->advectedTime;
  return 0;
}
int lsg_advect_get_current_time_step(lsg_advect *a, double *dt) {
  *dt = a->currentTimeStep;
  return 0;
}
int lsg_advect_get_number_of_time_steps(lsg_advect *a, unsigned *n) {
  *n = a->numSteps;
  return 0;
}
int lsg_advect_get_point_updates(lsg_advect *a, double *n) {
  *n = a->pointUpdates;
  return 0;
}
int lsg_advect_set_profiling(lsg_advect *a, int on) {
  a->dropGraphs();
  if (int e = ctxInit()) return e;
  double z = 0;
  LSG_HIP(hipMemcpy(&g_ctx.sc->rateBytes, &z, sizeof(double), hipMemcpyHostToDevice));
  LSG_HIP(hipMemcpy(&g_ctx.sc->starBytes, &z, sizeof(double), hipMemcpyHostToDevice));
  LSG_HIP(hipMemcpy(&g_ctx.sc->star2Bytes, &z, sizeof(double), hipMemcpyHostToDevice));
  a->profiling = on != 0;
  a->rateMs = 0;
  a->rateLaunches = 0;
  a->rateBytes = 0;
  a->starMs = 0;
  a->starLaunches = 0;
  a->starBytes = 0;
  a->star2Ms = 0;
  a->star2Launches = 0;
  a->star2Bytes = 0;
  return 0;
}
int lsg_advect_get_rate_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes) {
  *ms = a->rateMs;
  *launches = a->rateLaunches;
  *bytes = a->rateBytes;
  return 0;
}
int lsg_advect_get_star_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes) {
  *ms = a->starMs;
  *launches = a->starLaunches;
  *bytes = a->starBytes;
  return 0;
}
int lsg_advect_get_expand2_kernel_stats(lsg_advect *a, double *ms, int64_t *launches, double *bytes) {
  *ms = a->star2Ms;
  *launches = a->star2Launches;
  *bytes = a->star2Bytes;
  return 0;
}

int lsg_weno_eval(int order, int64_t n, const double *x, double dx, int plus, double *out) {
  if (order != 3 && order != 5) return fail(-1, "WENO order must be 3 or 5.");
  if (int e = ctxInit()) return e;
  const int w = order == 3 ? 5 : 7;
  double *dx_ = nullptr, *dout = nullptr;
  LSG_HIP(hipMalloc(&dx_, (size_t)n * w * sizeof(double) + 8));
  LSG_HIP(hipMalloc(&dout, (size_t)n * sizeof(double) + 8));
  LSG_HIP(hipMemcpyAsync(dx_, x, (size_t)n * w * sizeof(double), hipMemcpyHostToDevice, g_ctx.s));
  launchWeno(g_ctx.s, order, (uint32_t)n, dx_, dx, plus, dout);
  LSG_HIP(hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  (void)hipFree(dx_);
  (void)hipFree(dout);
  return 0;
}

// ---- learned weno5: network training ----------------------------------------

struct lsg_mlp {
  float *P = nullptr, *M = nullptr, *V = nullptr, *G = nullptr, *partials = nullptr;
  float *X = nullptr, *T = nullptr;
  int64_t n = 0, cap = 0;
  int rows = 0;
  int step = 0;
  ~lsg_mlp() {
    for (float *p : {P, M, V, G, partials, X, T})
      if (p) (void)hipFree(p);
  }
};

int lsg_mlp_create(lsg_mlp **out, const float *params) {
  using namespace lsg;
  if (!out || !params) return fail(-1, "lsg_mlp_create: null argument");
  if (int e = ctxInit()) return e;
  auto *m = new lsg_mlp();
  m->rows = 1024;
  const size_t pb = MLP_PSTRIDE * sizeof(float);
  if (hipMalloc(&m->P, pb) != hipSuccess || hipMalloc(&m->M, pb) != hipSuccess || hipMalloc(&m->V, pb) != hipSuccess ||
      hipMalloc(&m->G, pb) != hipSuccess || hipMalloc(&m->partials, pb * m->rows) != hipSuccess) {
    delete m;
    return fail(-20, "lsg_mlp_create: out of device memory");
  }
  LSG_HIP(hipMemcpy(m->P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  LSG_HIP(hipMemset(m->M, 0, pb));
  LSG_HIP(hipMemset(m->V, 0, pb));
  *out = m;
  return 0;
}

int lsg_mlp_destroy(lsg_mlp *m) {
  delete m;
  return 0;
}

int lsg_mlp_set_samples(lsg_mlp *m, int64_t n, const float *X, const float *T) {
  if (n <= 0 || n > (int64_t)UINT32_MAX / 8) return fail(-1, "lsg_mlp_set_samples: sample count out of range");
  if (n > m->cap) {
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    m->X = m->T = nullptr;
    m->cap = 0;
    LSG_HIP(hipMalloc(&m->X, (size_t)n * 8 * sizeof(float)));
    LSG_HIP(hipMalloc(&m->T, (size_t)n * sizeof(float)));
    m->cap = n;
  }
  LSG_HIP(hipMemcpy(m->X, X, (size_t)n * 8 * sizeof(float), hipMemcpyHostToDevice));
  LSG_HIP(hipMemcpy(m->T, T, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  m->n = n;
  return 0;
}

static int mlpRows(int64_t count, int maxRows) {
  int64_t r = (count + 63) / 64;
  if (r > maxRows) r = maxRows;
  return r < 1 ? 1 : (int)r;
}

int lsg_mlp_loss_grad(lsg_mlp *m, int64_t offset, int64_t count, double *loss, float *grad) {
  using namespace lsg;
  if (offset < 0 || count <= 0 || offset + count > m->n) return fail(-1, "lsg_mlp_loss_grad: sample range out of bounds");
  launchMlpGrad(g_ctx.s, m->X + offset * 8, m->T + offset, (uint32_t)count, m->P, m->partials,
                mlpRows(count, m->rows), m->G);
  float h[MLP_PSTRIDE];
  LSG_HIP(hipMemcpyAsync(h, m->G, sizeof(h), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  if (grad) memcpy(grad, h, MLP_NPARAM * sizeof(float));
  if (loss) *loss = (double)h[MLP_NPARAM] / (double)count;
  return 0;
}

int lsg_mlp_train(lsg_mlp *m, int iters, int64_t batch, double lr, double *lastLoss) {
  using namespace lsg;
  if (m->n <= 0) return fail(-1, "lsg_mlp_train: no samples");
  if (batch <= 0 || batch > m->n) batch = m->n;
  int64_t off = 0;
  for (int it = 0; it < iters; ++it) {
    if (off + batch > m->n) off = 0;
    launchMlpGrad(g_ctx.s, m->X + off * 8, m->T + off, (uint32_t)batch, m->P, m->partials, mlpRows(batch, m->rows),
                  m->G);
    launchMlpAdam(g_ctx.s, m->P, m->M, m->V, m->G, ++m->step, (float)lr);
    off += batch;
  }
  if (lastLoss) {
    float l;
    LSG_HIP(hipMemcpyAsync(&l, m->G + MLP_NPARAM, sizeof(float), hipMemcpyDeviceToHost, g_ctx.s));
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    *lastLoss = (double)l / (double)batch;
  }
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

int lsg_mlp_get_params(lsg_mlp *m, float *out) {
  LSG_HIP(hipMemcpy(out, m->P, lsg::MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

int lsg_weno5_learned_eval(int64_t n, const double *x, double dx, int plus, const float *params, double *out) {
  using namespace lsg;
  if (n <= 0 || n > (int64_t)UINT32_MAX) return fail(-1, "lsg_weno5_learned_eval: n out of range");
  if (int e = ctxInit()) return e;
  std::vector<uint16_t> b(MLP_NPARAM);
  for (int i = 0; i < MLP_NPARAM; ++i) b[i] = bf16Bits(params[i]);
  double *dx_ = nullptr, *dout = nullptr;
  uint16_t *dp = nullptr;
  LSG_HIP(hipMalloc(&dx_, (size_t)n * 7 * sizeof(double)));
  LSG_HIP(hipMalloc(&dout, (size_t)n * sizeof(double)));
  LSG_HIP(hipMalloc(&dp, MLP_NPARAM * sizeof(uint16_t)));
  LSG_HIP(hipMemcpyAsync(dx_, x, (size_t)n * 7 * sizeof(double), hipMemcpyHostToDevice, g_ctx.s));
  LSG_HIP(hipMemcpyAsync(dp, b.data(), MLP_NPARAM * sizeof(uint16_t), hipMemcpyHostToDevice, g_ctx.s));
  launchWenoLearned(g_ctx.s, (uint32_t)n, dx_, dx, plus, dp, dout);
  LSG_HIP(hipMemcpyAsync(out, dout, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, g_ctx.s));
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  (void)hipFree(dx_);
  (void)hipFree(dout);
  (void)hipFree(dp);
  return 0;
}

// ---- batch -------------------------------------------------------------------

int lsg_batch_create(lsg_batch **out, lsg_domain *const *domains, int n) {
  if (!out || !domains || n <= 0) return fail(-1, "lsg_batch_create: no level sets");
  if (n > 65535) return fail(-1, "lsg_batch_create: at most 65535 level sets");
  if (int e = ctxInit()) return e;
  auto *b = new lsg_batch();
  b->D = domains[0]->D;
  for (int k = 0; k < n; ++k) {
    if (!domains[k] || domains[k]->D != b->D) {
      delete b;
      return fail(-1, "lsg_batch_create: level sets must exist and have the same dimension");
    }
    b->doms.push_back(domains[k]);
  }
  b->segsH.assign(n, SegInfo{});
  b->velSet.assign(n, 0);
  *out = b;
  return 0;
}

int lsg_batch_destroy(lsg_batch *b) {
  delete b;
  return 0;
}

int lsg_batch_set_velocity_builtin(lsg_batch *b, int k, int kind, int dir, int kk, double b0, double aPos, double aNeg,
                                   const double *vec3) {
  if (k < 0 || k >= (int)b->doms.size()) return fail(-1, "batch: level set index out of range");
  if (kind < 0 || kind > 2) return fail(-1, "unknown velocity kind");
  if (dir < 0 || dir > 2) return fail(-1, "velocity direction must be 0, 1 or 2");
  Vel &v = b->segsH[k].vel;
  v.kind = kind;
  v.dir = dir;
  v.k = kk;
  v.b = b0;
  v.aPos = aPos;
  v.aNeg = aNeg;
  for (int i = 0; i < 3; ++i) v.vec[i] = vec3 ? vec3[i] : 0.0;
  b->velSet[k] = 1;
  if (b->assembled) {
    b->dropGraphs();
    LSG_HIP(hipMemcpy(&b->segsD[k], &b->segsH[k], sizeof(SegInfo), hipMemcpyHostToDevice));
  }
  return 0;
}

int lsg_batch_set_schemes(lsg_batch *b, int spatial, int temporal, double cfl) {
  if (spatial < 0 || spatial > 12) return fail(-1, "Advect: Discretization scheme not found.");
  const bool reassemble = b->assembled && ((temporal != 0) != (b->temporal != 0));
  b->spatial = spatial;
  b->temporal = (temporal == 1 || temporal == 2) ? temporal : 0;
  b->cfl = cfl;
  b->dropGraphs();
  if (reassemble) {
    if (int e = batchWriteBack(b)) return e;
    b->assembled = false;
  }
  return 0;
}

int lsg_batch_apply(lsg_batch *b) {
  if (g_inVelocityUpdate) return fail(-1, "batch advection called from a velocity update callback");
  if (int e = batchPrepare(b)) return e;
  return batchStep(b);
}

int lsg_batch_get_advected_times(lsg_batch *b, double *t) {
  const size_t S = b->doms.size();
  if (!b->segAdvTime) {
    for (size_t k = 0; k < S; ++k) t[k] = 0;
    return 0;
  }
  LSG_HIP(hipMemcpy(t, b->segAdvTime, S * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int lsg_batch_get_point_updates(lsg_batch *b, double *n, unsigned *steps) {
  if (n) *n = b->pointUpdates;
  if (steps) *steps = b->steps;
  return 0;
}

int lsg_batch_write_back(lsg_batch *b) { return batchWriteBack(b); }

int lsg_set_single_pass_band_ops(int on) {
  if (int e = ctxInit()) return e;
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  setSinglePassBandOps(on != 0);
  return 0;
}
int lsg_batch_set_scan_blocks(lsg_batch *b, int blocks) {
  if (blocks < 8 || blocks > SCAN_BLOCKS_MAX) return fail(-1, "batch: scan grid out of range");
  b->scanBlocks = blocks;
  b->dropGraphs();
  return 0;
}

int lsg_batch_autotune(lsg_batch *b, int *chosen, double *ms) {
  // time one step per candidate scan grid, each from the same state (a copy
  // of every level set, restored before each trial and at the end)
  if (int e = batchPrepare(b)) return e;
  if (int e = batchWriteBack(b)) return e;
  const size_t S = b->doms.size();
  std::vector<double> adv(S);
  LSG_HIP(hipMemcpy(adv.data(), b->segAdvTime, S * sizeof(double), hipMemcpyDeviceToHost));
  std::vector<lsg_domain *> snap(S);
  auto cleanup = [&]() {
    for (lsg_domain *d : snap) delete d;
  };
  for (size_t k = 0; k < S; ++k) {
    snap[k] = new lsg_domain();
    if (int e = copyDomain(snap[k], b->doms[k])) {
      cleanup();
      return e;
    }
  }
  auto restore = [&]() -> int {
    for (size_t k = 0; k < S; ++k)
      if (int e = copyDomain(b->doms[k], snap[k])) return e;
    LSG_HIP(hipStreamSynchronize(g_ctx.s));
    b->assembled = false;
    return batchPrepare(b);
  };
  const double pu0 = b->pointUpdates;
  const unsigned st0 = b->steps;
  const bool graphs = b->useGraphs;
  b->useGraphs = false;
  const int cand[] = {256, 512, 1024, 2048, 4096};
  int best = b->scanBlocks;
  double bestMs = 1e300;
  int err = 0;
  for (int c : cand) {
    if ((err = restore())) break;
    b->scanBlocks = c;
    // two steps from the restored state; the second one is timed (the
    // first also touches freshly restored memory and grows buffers)
    if ((err = batchStep(b))) break;
    LSG_HIP(hipEventRecord(g_ctx.ev0, g_ctx.s));
    if ((err = batchStep(b))) break;
    LSG_HIP(hipEventRecord(g_ctx.ev1, g_ctx.s));
    LSG_HIP(hipEventSynchronize(g_ctx.ev1));
    float m = 0;
    LSG_HIP(hipEventElapsedTime(&m, g_ctx.ev0, g_ctx.ev1));
    if (m < bestMs) {
      bestMs = m;
      best = c;
    }
  }
  if (!err) err = restore();
  cleanup();
  b->useGraphs = graphs;
  if (err) return err;
  LSG_HIP(hipMemcpy(b->segAdvTime, adv.data(), S * sizeof(double), hipMemcpyHostToDevice));
  b->pointUpdates = pu0;
  b->steps = st0;
  b->scanBlocks = best;
  b->dropGraphs();
  if (chosen) *chosen = best;
  if (ms) *ms = bestMs;
  return 0;
}

int lsg_device_synchronize(void) {
  if (int e = ctxInit()) return e;
  LSG_HIP(hipStreamSynchronize(g_ctx.s));
  return 0;
}

}  // extern "C"
