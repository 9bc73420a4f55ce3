// Host build of the GPU per-item logic (viennals_amd/csrc/lsg_items.hpp)
// driven through the same frame/bitmap pipeline as the HIP kernels, with
// serial scans.  TEST INFRASTRUCTURE ONLY: it lets the CPU test suite check
// the device algorithm (candidate bitmaps, rank lookups, compaction, index
// rebuild, schemes) against the CPU oracle and catch out-of-range indexing
// before any GPU run.  The product library never uses it.
#ifndef LSG_HOST_ONLY
#define LSG_HOST_ONLY
#endif
#include "../../viennals_amd/csrc/lsg_geometry.hpp"
#include "../../viennals_amd/csrc/lsg_items.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

using namespace lsg;

namespace {

constexpr int MARGIN = 24;
constexpr int REFRAME_SLACK = 12;


struct EBand {
  int D = 3;
  GridInfo g{};
  Frame f{};
  int width = 1;
  std::vector<uint64_t> bits;
  std::vector<uint32_t> pre;
  std::vector<double> val;
  std::vector<uint32_t> pos;
  BandView view() const {
    BandView v;
    v.bits = bits.data();
    v.pre = pre.data();
    v.val = val.data();
    v.pos = pos.data();
    v.f = f;
    v.g = g;
    v.rb = (uint32_t)f.base;
    return v;
  }
};

void check(bool c, const char *m) {
  if (!c) throw std::runtime_error(m);
}

Frame makeFrame(const EBand &d, const int32_t *bbmin, const int32_t *bbmax) {
  Frame f{};
  for (int k = 0; k < 3; ++k) {
    if (k >= d.D) {
      f.o[k] = 0;
      f.ext[k] = 1;
      continue;
    }
    if (d.g.bounded(k)) {
      f.o[k] = d.g.lo[k];
      f.ext[k] = d.g.hi[k] - d.g.lo[k] + 1;
    } else {
      int64_t lo = std::max<int64_t>((int64_t)bbmin[k] - MARGIN, d.g.lo[k]);
      int64_t hi = std::min<int64_t>((int64_t)bbmax[k] + MARGIN, d.g.hi[k]);
      f.o[k] = (int32_t)lo;
      f.ext[k] = (int32_t)(hi - lo + 1);
    }
  }
  const int64_t bitsTotal = (int64_t)((f.ext[0] + 63) >> 6) * 64 * f.ext[1] * f.ext[2];
  check(bitsTotal < ((int64_t)1 << 32) - 64, "frame too large");
  f.finish();
  bool inFrame = true;
  for (int k = 0; k < d.D; ++k) inFrame = inFrame && f.o[k] == d.g.lo[k];
  f.base = inFrame ? 0 : 1;
  return f;
}

void upload(EBand &d, const std::vector<int32_t> &idx, const std::vector<double> &val, int width,
            const Frame *forced = nullptr) {
  const int D = d.D;
  const size_t n = val.size();
  int32_t bbmin[3] = {0, 0, 0}, bbmax[3] = {0, 0, 0};
  bool any = false;
  for (size_t i = 1; i < n; ++i) {
    for (int k = 0; k < D; ++k) {
      int32_t c = idx[i * D + k];
      if (!any || c < bbmin[k]) bbmin[k] = c;
      if (!any || c > bbmax[k]) bbmax[k] = c;
    }
    any = true;
  }
  if (!any)
    for (int k = 0; k < D; ++k) bbmin[k] = bbmax[k] = (d.g.bounded(k) ? d.g.lo[k] : 0);
  d.f = forced ? *forced : makeFrame(d, bbmin, bbmax);
  const Frame &f = d.f;
  d.pos.assign(n, 0xFFFFFFFFu);
  d.bits.assign(f.nwords + 1, 0);
  for (size_t i = 0; i < n; ++i) {
    if (i == 0 && f.base) continue;
    int32_t r[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) r[k] = idx[i * D + k] - f.o[k];
    check(r[0] >= 0 && r[0] < f.ext[0] && r[1] >= 0 && r[1] < f.ext[1] && r[2] >= 0 && r[2] < f.ext[2],
          "entry outside frame");
    uint32_t p = f.encode(r[0], r[1], r[2]);
    check(i <= (size_t)f.base || p > d.pos[i - 1], "runs not sorted");
    d.pos[i] = p;
    d.bits[p >> 6] |= 1ull << (p & 63);
  }
  d.pre.assign(f.nwords + 1, 0);
  uint32_t acc = 0;
  for (int64_t w = 0; w <= f.nwords; ++w) {
    d.pre[w] = acc;
    if (w < f.nwords) acc += (uint32_t)__builtin_popcountll(d.bits[w]);
  }
  d.val = val;
  d.width = width;
}

void download(const EBand &d, std::vector<int32_t> &idx, std::vector<double> &val) {
  const int D = d.D;
  const size_t n = d.val.size();
  idx.assign(n * D, 0);
  val = d.val;
  for (size_t i = 0; i < n; ++i) {
    if (i == 0 && d.f.base) {
      for (int k = 0; k < D; ++k) idx[k] = d.g.lo[k];
      continue;
    }
    int32_t r[3];
    d.f.decode(d.pos[i], r[0], r[1], r[2]);
    for (int k = 0; k < D; ++k) idx[i * D + k] = r[k] + d.f.o[k];
  }
}

// compaction of candidates (cpos, out) into a new band with the same frame
EBand compact(const EBand &in, const std::vector<uint32_t> &cpos, const std::vector<uint32_t> &cpre,
              const std::vector<double> &out) {
  const uint32_t nc = (uint32_t)out.size();
  std::vector<uint32_t> ks(nc + 1);
  uint32_t acc = 0;
  for (uint32_t i = 0; i < nc; ++i) {
    ks[i] = acc;
    acc += keepFlag(out.data(), i);
  }
  ks[nc] = acc;
  EBand o;
  o.D = in.D;
  o.g = in.g;
  o.f = in.f;
  o.val.resize(acc);
  o.pos.resize(acc);
  for (uint32_t i = 0; i < nc; ++i)
    if (ks[i + 1] != ks[i]) {
      o.val[ks[i]] = out[i];
      o.pos[ks[i]] = cpos[i];
    }
  o.bits.assign(in.f.nwords + 1, 0);
  o.pre.assign(in.f.nwords + 1, 0);
  for (int64_t w = 0; w <= in.f.nwords; ++w) {
    check((size_t)in.f.base + cpre[w] <= nc, "cpre beyond candidates");
    buildIndexWord(cpre.data(), ks.data(), cpos.data(), w, in.f.nwords, in.f.base, nc, o.bits[w], o.pre[w]);
  }
  return o;
}

void candidates(const EBand &in, const std::vector<uint64_t> &C, std::vector<uint32_t> &cpos, std::vector<uint32_t> &cpre) {
  const int64_t nw = in.f.nwords;
  cpre.assign(nw + 1, 0);
  uint32_t acc = 0;
  for (int64_t w = 0; w < nw; ++w) {
    cpre[w] = acc;
    acc += (uint32_t)__builtin_popcountll(C[w]);
  }
  cpre[nw] = acc;
  cpos.assign(acc + in.f.base, 0);
  if (in.f.base) cpos[0] = 0xFFFFFFFFu;
  for (int64_t w = 0; w < nw; ++w) {
    uint64_t c = C[w];
    size_t t = in.f.base + cpre[w];
    while (c) {
      int b = __builtin_ctzll(c);
      cpos[t++] = (uint32_t)((w << 6) | b);
      c &= c - 1;
    }
  }
}

// The single-pass band operation of the HIP kernels (kBandOp), tile by
// tile in order: candidate words of BAND_TILE_WORDS words, the rule at every
// slot (tileSlot), keep flags against the previous slot, compaction, bitmap
// words and prefix counts.  Used instead of the candidate-list pipeline
// when g_tiled.
constexpr int TW = 32;
bool g_tiled = true;

enum { MODE_STAR = 0, MODE_UNION = 1, MODE_SELF = 2 };

template <class At, int D>
EBand tiledOp(const EBand &in, const At &at, int mode, const std::vector<uint64_t> *B = nullptr) {
  const Frame &f = in.f;
  const int64_t nw = f.nwords;
  const uint32_t base = (uint32_t)f.base;
  EBand o;
  o.D = in.D;
  o.g = in.g;
  o.f = f;
  o.bits.assign(nw + 1, 0);
  o.pre.assign(nw + 1, 0);
  o.val.assign(base, 0.0);
  o.pos.assign(base, 0xFFFFFFFFu);
  const uint32_t nTiles = (uint32_t)((nw + TW - 1) / TW);
  uint32_t P = 0;
  for (uint32_t t = 0; t < nTiles; ++t) {
    const uint32_t w0 = t * TW;
    uint64_t cw[TW], kw[TW];
    uint32_t cpre[TW + 1], kpre[TW + 1];
    uint32_t acc = 0;
    for (int j = 0; j < TW; ++j) {
      const int64_t w = (int64_t)w0 + j;
      uint64_t C = 0;
      if (w < nw) {
        if (mode == MODE_STAR) C = candStarWord<D>(in.bits.data(), f, in.g, w);
        else if (mode == MODE_UNION) C = in.bits[w] | (*B)[w];
        else C = in.bits[w];
      }
      cw[j] = C;
      kw[j] = 0;
      cpre[j] = acc;
      acc += (uint32_t)__builtin_popcountll(C);
    }
    cpre[TW] = acc;
    const uint32_t nc = acc;
    const bool hasPred = t > 0 || base;
    const uint32_t nslots = (nc > 0 || (t == 0 && base)) ? nc + 1 : 0;
    std::vector<double> stage;
    double prev = 0.0;
    for (uint32_t s = 0; s < nslots; ++s) {
      uint32_t p;
      bool eval, cand;
      int j, b;
      tileSlot<TW>(f, t, w0, cw, cpre, nc, s, p, eval, cand, j, b);
      check(!eval || p == 0xFFFFFFFFu || (int64_t)(p >> 6) < nw, "slot position outside the frame");
      const double v = eval ? at(p) : 0.0;
      if (s == 0 && t == 0 && base) o.val[0] = v;
      const bool keep = cand && ((s == 1 && !hasPred) || !isUndef(v) || v != prev);
      if (keep) {
        stage.push_back(v);
        kw[j] |= 1ull << b;
      }
      prev = v;
    }
    acc = 0;
    for (int j = 0; j < TW; ++j) {
      kpre[j] = acc;
      acc += (uint32_t)__builtin_popcountll(kw[j]);
    }
    kpre[TW] = acc;
    check(acc == stage.size(), "keep words and kept values disagree");
    for (uint32_t k = 0; k < acc; ++k) {
      const int jj = wordOfItem<TW>(kpre, k);
      const int bb = selectBit(kw[jj], (int)(k - kpre[jj]));
      o.val.push_back(stage[k]);
      o.pos.push_back(((w0 + (uint32_t)jj) << 6) | (uint32_t)bb);
    }
    for (int j = 0; j < TW; ++j) {
      const int64_t w = (int64_t)w0 + j;
      if (w < nw) {
        o.bits[w] = kw[j];
        o.pre[w] = P + kpre[j];
      }
    }
    P += acc;
  }
  o.bits[nw] = 0;
  o.pre[nw] = P;
  return o;
}

template <int D> EBand starOp(const EBand &in, const OpParams &op) {
  if (g_tiled) return tiledOp<AtStar<D>, D>(in, AtStar<D>{in.view(), op}, MODE_STAR);
  const int64_t nw = in.f.nwords;
  std::vector<uint64_t> C(nw + 1, 0);
  for (int64_t w = 0; w < nw; ++w) C[w] = candStarWord<D>(in.bits.data(), in.f, in.g, w);
  std::vector<uint32_t> cpos, cpre;
  candidates(in, C, cpos, cpre);
  BandView v = in.view();
  std::vector<double> out(cpos.size());
  for (uint32_t i = 0; i < cpos.size(); ++i) {
    int32_t c[3];
    candCoords(v, i, cpos[i], c);
    double center, nb[2 * D];
    gatherStar<D>(v, c, center, nb);
    out[i] = (op.kind == OP_EXPAND) ? ruleExpand<D>(center, nb, op.a, op.b) : ruleRebuild<D>(center, nb, op.a);
  }
  return compact(in, cpos, cpre, out);
}

// Two Expand cycles in one operation (the candidate-list pipeline's
// OP_EXPAND2): candidates from the star dilation of the star dilation.
bool g_pairs = false;
template <int D> EBand starOp2(const EBand &in, const OpParams &op) {
  const int64_t nw = in.f.nwords;
  std::vector<uint64_t> T(nw + 1, 0), C(nw + 1, 0);
  for (int64_t w = 0; w < nw; ++w) T[w] = candStarWord<D>(in.bits.data(), in.f, in.g, w);
  for (int64_t w = 0; w < nw; ++w) C[w] = candStarWord<D>(T.data(), in.f, in.g, w);
  std::vector<uint32_t> cpos, cpre;
  candidates(in, C, cpos, cpre);
  BandView v = in.view();
  std::vector<double> out(cpos.size());
  for (uint32_t i = 0; i < cpos.size(); ++i) {
    int32_t c[3];
    candCoords(v, i, cpos[i], c);
    out[i] = expandTwice<D>(v, c, op.a, op.a2, op.b);
  }
  return compact(in, cpos, cpre, out);
}

// updateLevelSet + rebuildLS as one star operation (OP_UPDATE_REBUILD).
template <int D> EBand updRebuildOp(const EBand &in, const std::vector<double> &rates, double dt, double cutoff) {
  const int64_t nw = in.f.nwords;
  std::vector<uint64_t> C(nw + 1, 0);
  for (int64_t w = 0; w < nw; ++w) C[w] = candStarWord<D>(in.bits.data(), in.f, in.g, w);
  std::vector<uint32_t> cpos, cpre;
  candidates(in, C, cpos, cpre);
  BandView v = in.view();
  std::vector<double> out(cpos.size());
  for (uint32_t i = 0; i < cpos.size(); ++i) {
    int32_t c[3];
    candCoords(v, i, cpos[i], c);
    out[i] = updateRebuild<D>(v, c, rates.data(), dt, cutoff);
  }
  return compact(in, cpos, cpre, out);
}

template <int D> EBand selfOp(const EBand &in, const OpParams &op, const std::vector<double> *rates, double dt) {
  if (g_tiled) {
    OpParams o2 = op;
    o2.rates = rates ? rates->data() : nullptr;
    o2.dt = &dt;
    return tiledOp<AtSelf<D>, D>(in, AtSelf<D>{in.view(), o2}, MODE_SELF);
  }
  BandView v = in.view();
  std::vector<double> out(in.val.size());
  for (uint32_t i = 0; i < in.val.size(); ++i) {
    double c = in.val[i];
    if (op.kind == OP_REDUCE) out[i] = ruleReduce(c, op.a);
    else if (op.kind == OP_UPDATE) out[i] = ruleUpdate(c, dt, (*rates)[i]);
    else {
      int32_t cc[3];
      candCoords(v, i, in.pos[i], cc);
      double center, nb[2 * D];
      gatherStar<D>(v, cc, center, nb);
      out[i] = rulePrune<D>(c, nb, op.kind == OP_PRUNE_RSZ);
    }
  }
  return compact(in, in.pos, in.pre, out);
}

EBand booleanOp(const EBand &A, const EBand &B, int kind, double wT, double wS) {
  check(std::memcmp(&A.f, &B.f, sizeof(Frame)) == 0, "boolean on different frames");
  const int64_t nw = A.f.nwords;
  std::vector<uint64_t> C(nw + 1, 0);
  for (int64_t w = 0; w < nw; ++w) C[w] = A.bits[w] | B.bits[w];
  std::vector<uint32_t> cpos, cpre;
  candidates(A, C, cpos, cpre);
  BandView va = A.view(), vb = B.view();
  if (g_tiled) {
    EBand o = (A.D == 2) ? tiledOp<AtBool, 2>(A, AtBool{va, vb, kind, wT, wS, nullptr}, MODE_UNION, &B.bits)
                         : tiledOp<AtBool, 3>(A, AtBool{va, vb, kind, wT, wS, nullptr}, MODE_UNION, &B.bits);
    o.width = A.width;
    return o;
  }
  std::vector<double> out(cpos.size());
  for (uint32_t i = 0; i < cpos.size(); ++i) {
    int32_t c[3];
    candCoords(va, i, cpos[i], c);
    out[i] = ruleBoolean(kind, va.value(c[0], c[1], c[2]), vb.value(c[0], c[1], c[2]), wT, wS);
  }
  EBand o = compact(A, cpos, cpre, out);
  o.width = A.width;
  return o;
}

// ---- ops mirroring lsg_host ----

template <int D> void doExpand(EBand &d, int width) {
  if (width <= d.width) return;
  const double totalLimit = width * 0.5;
  const int startWidth = d.width;
  const int cycles = width - startWidth;
  for (int cyc = 0; cyc < cycles;) {
    OpParams op{};
    op.a = (startWidth + cyc + 1) * 0.5;
    op.b = totalLimit;
    if (g_pairs && !g_tiled && cycles - cyc >= 2) {
      op.kind = OP_EXPAND2;
      op.a2 = (startWidth + cyc + 2) * 0.5;
      d = starOp2<D>(d, op);
      cyc += 2;
    } else {
      op.kind = OP_EXPAND;
      d = starOp<D>(d, op);
      cyc += 1;
    }
  }
  d.width = width;
}
template <int D> void doReduce(EBand &d, int width) {
  if (width >= d.width) return;
  OpParams op{};
  op.kind = OP_REDUCE;
  op.a = width * 0.5;
  d = selfOp<D>(d, op, nullptr, 0);
  d.width = width;
}
template <int D> void doPrune(EBand &d, bool rsz) {
  OpParams op{};
  op.kind = rsz ? OP_PRUNE_RSZ : OP_PRUNE;
  d = selfOp<D>(d, op, nullptr, 0);
  d.width = 2;
}
template <int D> void doRebuild(EBand &d, bool sllf = false) {
  OpParams op{};
  op.kind = OP_REBUILD;
  op.a = sllf ? 1.5 : 1.0;
  d = starOp<D>(d, op);
  d.width = sllf ? 3 : 2;
}
template <int D> void doBoolean(EBand &a, const EBand &b, int kind, double wT, double wS) {
  a = booleanOp(a, b, kind, wT, wS);
  doPrune<D>(a, true);
  doPrune<D>(a, false);
}

void bbox(const EBand &d, int32_t mn[3], int32_t mx[3], bool &any) {
  any = false;
  for (size_t i = 0; i < d.val.size(); ++i) {
    if (i == 0 && d.f.base) continue;
    int32_t r[3];
    d.f.decode(d.pos[i], r[0], r[1], r[2]);
    for (int k = 0; k < 3; ++k) {
      int32_t a = r[k] + d.f.o[k];
      if (!any || a < mn[k]) mn[k] = a;
      if (!any || a > mx[k]) mx[k] = a;
    }
    any = true;
  }
}

void reframe(EBand &d, bool force) {
  int32_t mn[3], mx[3];
  bool any;
  bbox(d, mn, mx, any);
  if (!any) return;
  bool need = force;
  for (int k = 0; k < d.D; ++k) {
    if (d.g.bounded(k)) continue;
    if (mn[k] - d.f.o[k] < REFRAME_SLACK || (d.f.o[k] + d.f.ext[k] - 1) - mx[k] < REFRAME_SLACK) need = true;
  }
  if (!need) return;
  std::vector<int32_t> idx;
  std::vector<double> val;
  download(d, idx, val);
  upload(d, idx, val, d.width);
}

void commonFrame(EBand &a, EBand &b) {
  if (std::memcmp(&a.f, &b.f, sizeof(Frame)) == 0) return;
  int32_t amn[3], amx[3], bmn[3], bmx[3], mn[3], mx[3];
  bool aa, ba;
  bbox(a, amn, amx, aa);
  bbox(b, bmn, bmx, ba);
  for (int k = 0; k < 3; ++k) {
    mn[k] = aa ? (ba ? std::min(amn[k], bmn[k]) : amn[k]) : (ba ? bmn[k] : 0);
    mx[k] = aa ? (ba ? std::max(amx[k], bmx[k]) : amx[k]) : (ba ? bmx[k] : 0);
  }
  Frame f = makeFrame(a, mn, mx);
  for (EBand *d : {&a, &b}) {
    std::vector<int32_t> idx;
    std::vector<double> val;
    download(*d, idx, val);
    upload(*d, idx, val, d->width, &f);
  }
}

int prepareWidth(int spatial) {
  static const int w[12] = {3, 5, 3, 5, 7, 3, 5, 7, 9, 8, 5, 7};
  return (spatial >= 0 && spatial < 12) ? w[spatial] : -1;
}
int schemeOrder(int s) {
  switch (s) {
  case 1: case 3: case 6: case 8: return 2;
  case 10: return 3;
  case 11: return 5;
  default: return 1;
  }
}

struct EAdvect {
  RateParams rp{};
  int temporal = 0;
  double cfl = 0.4999, dissipationAlpha = 1.0;
  double pointUpdates = 0;
  std::vector<double> rates;

  template <int D, int S> double ratesPass(EBand &d, double maxTimeStep, const double *ga) {
    BandView v = d.view();
    rates.assign(d.val.size(), 0.0);
    Vel vel = velOf(rp);
    double lmax = 0, la[3] = {0, 0, 0};
    for (uint32_t i = 0; i < d.val.size(); ++i) {
      double phi0 = d.val[i];
      if (isUndef(phi0) || fabs(phi0) > 0.5) continue;
      pointUpdates += 1;
      int32_t c[3];
      candCoords(v, i, d.pos[i], c);
      double velocity;
      // SLLF leaves la at the finalAlphas of the last point that set them
      rates[i] = ratePoint<D, S>(v, c, phi0, rp, vel, ga, la, velocity);
      if (velocity != 0.) {
        double a = fabs(velocity);
        if (lmax < a) lmax = a;
      }
    }
    const bool useAlphas = (S == S_LF || S == S_LLLF || S == S_LLF || S == S_LLFA || S == S_SLLF);
    const double *al = (S == S_LF) ? ga : la;
    return stageTimeStep(maxTimeStep, rp.cfl, lmax, D, al, useAlphas, d.g.dx);
  }

  template <int D> double stage(EBand &d, double maxTimeStep) {
    doExpand<D>(d, prepareWidth(rp.spatial));
    double ga[3] = {0, 0, 0};
    const int sid = schemeIdOf(rp.spatial);
    if (sid == S_LF) {
      BandView v = d.view();
      Vel vel = velOf(rp);
      for (uint32_t i = 0; i < d.val.size(); ++i) {
        double v0 = d.val[i];
        if (isUndef(v0) || fabs(v0) > 0.5) continue;
        int32_t c[3];
        candCoords(v, i, d.pos[i], c);
        globalAlphaPoint<D>(v, c, vel, ga);
      }
    }
    double dt;
    switch (sid) {
    case S_EO: dt = ratesPass<D, S_EO>(d, maxTimeStep, ga); break;
    case S_LF: dt = ratesPass<D, S_LF>(d, maxTimeStep, ga); break;
    case S_LLLF: dt = ratesPass<D, S_LLLF>(d, maxTimeStep, ga); break;
    case S_LLF: dt = ratesPass<D, S_LLF>(d, maxTimeStep, ga); break;
    case S_LLFA: dt = ratesPass<D, S_LLFA>(d, maxTimeStep, ga); break;
    case S_SLLF: dt = ratesPass<D, S_SLLF>(d, maxTimeStep, ga); break;
    case S_WENO3: dt = ratesPass<D, S_WENO3>(d, maxTimeStep, ga); break;
    default: dt = ratesPass<D, S_WENO5>(d, maxTimeStep, ga); break;
    }
    OpParams op{};
    if (g_pairs && !g_tiled) {  // the fused update and rebuild (OP_UPDATE_REBUILD)
      d = updRebuildOp<D>(d, rates, dt, rp.spatial == 9 ? 1.5 : 1.0);
      d.width = rp.spatial == 9 ? 3 : 2;
      return dt;
    }
    op.kind = OP_UPDATE;
    d = selfOp<D>(d, op, &rates, dt);
    d.width = 1;
    doRebuild<D>(d, rp.spatial == 9);
    return dt;
  }

  template <int D> void combine(EBand &d, EBand &init, double wT, double wS) {
    int steps = temporal == 1 ? 2 : (temporal == 2 ? 3 : 1);
    int ew = (int)std::ceil(2.0 * steps * cfl + 1);
    doExpand<D>(d, ew);
    doExpand<D>(init, ew);
    doBoolean<D>(d, init, B_COMBINE, wT, wS);
    doRebuild<D>(d, rp.spatial == 9);
  }

  template <int D> double step(EBand &d, double maxTimeStep) {
    double r;
    if (temporal == 0) {
      r = stage<D>(d, maxTimeStep);
    } else if (temporal == 1) {
      EBand init = d;
      double dt1 = stage<D>(d, maxTimeStep);
      double dt2 = stage<D>(d, dt1);
      combine<D>(d, init, 0.5, 0.5);
      r = 0.5 * dt1 + 0.5 * dt2;
    } else {
      EBand init = d;
      double dt1 = stage<D>(d, maxTimeStep);
      double dt2 = stage<D>(d, dt1);
      combine<D>(d, init, 0.75, 0.25);
      double dt3 = stage<D>(d, dt1);
      combine<D>(d, init, 1.0 / 3.0, 2.0 / 3.0);
      r = (dt1 + dt2 + 4.0 * dt3) / 6.0;
    }
    reframe(d, false);
    return r;
  }
};

}  // namespace

struct emu_domain {
  EBand b;
};

extern "C" {

emu_domain *emu_domain_new(int D, const int32_t *gmin, const int32_t *gmax, const int *bc, double delta) {
  auto *d = new emu_domain;
  d->b.D = D;
  for (int k = 0; k < 3; ++k) {
    d->b.g.lo[k] = d->b.g.hi[k] = 0;
    d->b.g.bc[k] = REFLECTIVE;
  }
  for (int k = 0; k < D; ++k) {
    d->b.g.bc[k] = bc[k];
    switch (bc[k]) {
    case INFINITE: d->b.g.lo[k] = -INF_IDX; d->b.g.hi[k] = INF_IDX; break;
    case PERIODIC: d->b.g.lo[k] = gmin[k]; d->b.g.hi[k] = gmax[k] - 1; break;
    case POS_INFINITE: d->b.g.lo[k] = gmin[k]; d->b.g.hi[k] = INF_IDX; break;
    case NEG_INFINITE: d->b.g.lo[k] = -INF_IDX; d->b.g.hi[k] = gmax[k]; break;
    default: d->b.g.lo[k] = gmin[k]; d->b.g.hi[k] = gmax[k]; break;
    }
  }
  d->b.g.D = D;
  d->b.g.dx = delta;
  std::vector<int32_t> idx(D);
  for (int k = 0; k < D; ++k) idx[k] = d->b.g.lo[k];
  upload(d->b, idx, std::vector<double>(1, POS), 1);
  return d;
}
void emu_domain_free(emu_domain *d) { delete d; }

int emu_domain_set(emu_domain *d, int64_t n, const int32_t *idx, const double *val, int width) {
  try {
    upload(d->b, std::vector<int32_t>(idx, idx + n * d->b.D), std::vector<double>(val, val + n), width);
  } catch (std::exception &e) {
    fprintf(stderr, "emu: %s\n", e.what());
    return -1;
  }
  return 0;
}
int64_t emu_domain_size(emu_domain *d) { return (int64_t)d->b.val.size(); }
int emu_domain_width(emu_domain *d) { return d->b.width; }
void emu_domain_get(emu_domain *d, int32_t *idx, double *val) {
  std::vector<int32_t> i;
  std::vector<double> v;
  download(d->b, i, v);
  std::memcpy(idx, i.data(), i.size() * sizeof(int32_t));
  std::memcpy(val, v.data(), v.size() * sizeof(double));
}

#define EMU_DISPATCH(stmt2, stmt3) \
  do {                             \
    try {                          \
      if (d->b.D == 2) stmt2;      \
      else stmt3;                  \
    } catch (std::exception & e) { \
      fprintf(stderr, "emu: %s\n", e.what()); \
      return -1;                   \
    }                              \
    return 0;                      \
  } while (0)

int emu_expand(emu_domain *d, int w) { EMU_DISPATCH(doExpand<2>(d->b, w), doExpand<3>(d->b, w)); }
int emu_reduce(emu_domain *d, int w) { EMU_DISPATCH(doReduce<2>(d->b, w), doReduce<3>(d->b, w)); }
int emu_prune(emu_domain *d, int rsz) { EMU_DISPATCH(doPrune<2>(d->b, rsz != 0), doPrune<3>(d->b, rsz != 0)); }
int emu_rebuild(emu_domain *d) { EMU_DISPATCH(doRebuild<2>(d->b), doRebuild<3>(d->b)); }
int emu_boolean(emu_domain *d, emu_domain *o, int kind, double wT, double wS) {
  EMU_DISPATCH((commonFrame(d->b, o->b), doBoolean<2>(d->b, o->b, kind, wT, wS)),
               (commonFrame(d->b, o->b), doBoolean<3>(d->b, o->b, kind, wT, wS)));
}

// Same parameter block as the oracle's lso_params.
struct emu_params {
  int spatial, temporal;
  double cfl, advectionTime;
  int singleStep;
  double dissipationAlpha;
  int calcNormals, checkDissipation;
  int velKind, velDir, velK;
  double velB, velAPos, velANeg;
  double vec[3];
};

int emu_advect(emu_domain *d, const emu_params *p, double *out) {
  EAdvect a;
  a.temporal = p->temporal;
  a.cfl = p->cfl;
  a.rp.spatial = p->spatial;
  a.rp.order = schemeOrder(p->spatial);
  a.rp.cfl = p->cfl;
  a.rp.alphaFactor = p->dissipationAlpha;
  a.rp.checkDissipation = p->checkDissipation;
  a.rp.sllfBaseEps = std::cbrt(std::numeric_limits<double>::epsilon());
  a.rp.calcNormals = (p->spatial == 10 || p->spatial == 11) ? (p->dissipationAlpha != 0.0) : p->calcNormals;
  a.rp.velKind = p->velKind;
  a.rp.velDir = p->velDir;
  a.rp.velK = p->velK;
  a.rp.velB = p->velB;
  a.rp.velAPos = p->velAPos;
  a.rp.velANeg = p->velANeg;
  for (int i = 0; i < 3; ++i) a.rp.vec[i] = p->vec[i];
  if (prepareWidth(p->spatial) < 0) return -2;
  double t = 0;
  unsigned steps = 0;
  try {
    if (p->advectionTime == 0.) {
      t = (d->b.D == 2) ? a.step<2>(d->b, std::numeric_limits<double>::max())
                        : a.step<3>(d->b, std::numeric_limits<double>::max());
      steps = 1;
    } else {
      while (t < p->advectionTime) {
        double dt = (d->b.D == 2) ? a.step<2>(d->b, p->advectionTime - t) : a.step<3>(d->b, p->advectionTime - t);
        if (!(dt > 0)) return -3;
        t += dt;
        ++steps;
        if (p->singleStep) break;
      }
    }
  } catch (std::exception &e) {
    fprintf(stderr, "emu: %s\n", e.what());
    return -1;
  }
  out[0] = t;
  out[1] = 0;
  out[2] = steps;
  out[3] = a.pointUpdates;
  return 0;
}

// The library's MakeGeometry plane (kind 0: a = origin, b = normal) or box
// (kind 1: a, b = corners; mask: setIgnoreBoundaryConditions) on a grid
// built like lsg_domain_create; returns the number of defined points written
// to idx / val (at most cap), or a negative error code.
int64_t emu_geometry(int kind, int D, const int32_t *gmin, const int32_t *gmax, const int *bc, double dx,
                     const double *a, const double *b, int mask, int32_t *idx, double *val, int64_t cap) {
  GridInfo g{};
  g.D = D;
  g.dx = dx;
  for (int k = 0; k < 3; ++k) g.bc[k] = REFLECTIVE;
  for (int k = 0; k < D; ++k) {
    g.bc[k] = bc[k];
    switch (bc[k]) {
    case INFINITE: g.lo[k] = -INF_IDX; g.hi[k] = INF_IDX; break;
    case PERIODIC: g.lo[k] = gmin[k]; g.hi[k] = gmax[k] - 1; break;
    case POS_INFINITE: g.lo[k] = gmin[k]; g.hi[k] = INF_IDX; break;
    case NEG_INFINITE: g.lo[k] = -INF_IDX; g.hi[k] = gmax[k]; break;
    default: g.lo[k] = gmin[k]; g.hi[k] = gmax[k]; break;
    }
  }
  std::vector<int32_t> pi;
  std::vector<double> pv;
  if (kind == 0) {
    const int e = lsg::planeBand(D, g, a, b, pi, pv);
    if (e) return e;
  } else {
    lsg::boxBand(D, g, a, b, pi, pv, mask);
  }
  const int64_t n = (int64_t)pv.size();
  for (int64_t i = 0; i < n && i < cap; ++i) {
    val[i] = pv[i];
    for (int k = 0; k < D; ++k) idx[i * D + k] = pi[i * D + k];
  }
  return n;
}

// 1: band operations through the single-pass tile pipeline; 0: candidate lists.
void emu_set_tiled(int on) { g_tiled = on != 0; }
void emu_set_expand_pairs(int on) { g_pairs = on != 0; }

// Bit helpers of the index kernels on the host.
void emu_deposit(const uint64_t *x, const uint64_t *m, uint64_t *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = lsg::depositBits(x[i], m[i]);
}
void emu_select(const uint64_t *w, const int32_t *k, int32_t *out, int64_t n) {
  for (int64_t i = 0; i < n; ++i) out[i] = lsg::selectBit(w[i], k[i]);
}

// Frame division helper on the host: q[i] = n[i] / d.
void emu_fastdiv(uint32_t d, const uint32_t *n, uint32_t *q, int64_t count) {
  lsg::FastDiv f;
  f.init(d);
  for (int64_t i = 0; i < count; ++i) q[i] = f.div(n[i]);
}

}  // extern "C"
