// HIP kernels of the level-set advection path, written for CDNA4 (gfx950,
// wave64).  Every kernel reads its problem size from device memory and runs
// a grid-stride loop over a fixed grid, so a whole time step is enqueued
// without host round trips.  The per-item logic lives in lsg_items.hpp.
//
// Once any kernel of a step flags a capacity overflow, every later kernel of
// the step returns immediately (the host then retries the step with larger
// buffers), so no kernel ever indexes past its buffers.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsg_device.hpp"
#include "lsg_items.hpp"
#include "lsg_kernels.hpp"

namespace lsg {

static constexpr int TPB = 256;
static int gridFor(size_t n, int cap = 4096) {
  size_t g = (n + TPB - 1) / TPB;
  if (g < 1) g = 1;
  if (g > (size_t)cap) g = cap;
  return (int)g;
}

BandView viewOf(const BandBufs &b, const Frame &f, const GridInfo &g) {
  BandView v;
  v.bits = b.bits;
  v.pre = b.pre;
  v.val = b.val;
  v.pos = b.pos;
  v.f = f;
  v.g = g;
  return v;
}

#define LSG_ABORT_IF_OVERFLOW(sc) \
  if ((sc)->overflow) return

// ------------------------------------------------------------- scans ----

// Block-wide exclusive scan of one value per thread (256 threads, 4 waves).
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t &total) {
  __shared__ uint32_t wsum[TPB / 64];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < TPB / 64; ++k) {
    uint32_t s = wsum[k];
    if (k < wid) woff += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return x - v + woff;
}

// Contiguous chunk [lo, hi) of block blockIdx.x for a device-side size n.
__device__ inline void blockChunk(uint32_t n, uint32_t &lo, uint32_t &hi) {
  const uint32_t tile = TPB * SCAN_ITEMS;
  uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  per = (per + tile - 1) / tile * tile;
  lo = blockIdx.x * per;
  hi = lo + per;
  if (lo > n) lo = n;
  if (hi > n) hi = n;
}

// Phase 1: per-block sums of Src::load(i).
template <class Src> __global__ void kScanReduce(Src src, const uint32_t *nptr, uint32_t *partials) {
  LSG_ABORT_IF_OVERFLOW(src.sc);
  const uint32_t n = src.size(nptr);
  uint32_t lo, hi;
  blockChunk(n, lo, hi);
  uint32_t acc = 0;
  for (uint32_t t = lo; t < hi; t += TPB * SCAN_ITEMS) {
    uint32_t i0 = t + threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      uint32_t i = i0 + k;
      if (i < hi) acc += src.load(i);
    }
  }
  uint32_t tot;
  blockExclusiveScan(acc, tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Phase 2: exclusive scan of the partials (SCAN_BLOCKS <= 2*TPB) in one block.
template <class Src> __global__ void kScanPartials(Src src, const uint32_t *nptr, uint32_t *partials, int nb) {
  LSG_ABORT_IF_OVERFLOW(src.sc);
  uint32_t v = (threadIdx.x < (unsigned)nb) ? partials[threadIdx.x] : 0;
  uint32_t v2 = (threadIdx.x + TPB < (unsigned)nb) ? partials[threadIdx.x + TPB] : 0;
  uint32_t t1, t2;
  uint32_t e1 = blockExclusiveScan(v, t1);
  uint32_t e2 = blockExclusiveScan(v2, t2);
  if (threadIdx.x < (unsigned)nb) partials[threadIdx.x] = e1;
  if (threadIdx.x + TPB < (unsigned)nb) partials[threadIdx.x + TPB] = e2 + t1;
  if (threadIdx.x == 0) src.total(nptr, t1 + t2);
}

// Phase 3: rescan each chunk with its offset; Src::store(i, prefix, value).
template <class Src> __global__ void kScanDown(Src src, const uint32_t *nptr, const uint32_t *partials) {
  LSG_ABORT_IF_OVERFLOW(src.sc);
  const uint32_t n = src.size(nptr);
  uint32_t lo, hi;
  blockChunk(n, lo, hi);
  uint32_t off = partials[blockIdx.x];
  for (uint32_t t = lo; t < hi; t += TPB * SCAN_ITEMS) {
    uint32_t i0 = t + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      uint32_t i = i0 + k;
      v[k] = (i < hi) ? src.load(i) : 0;
      s += v[k];
    }
    uint32_t tot;
    uint32_t ex = blockExclusiveScan(s, tot);
    uint32_t p = off + ex;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      uint32_t i = i0 + k;
      if (i < hi) src.store(i, p, v[k]);
      p += v[k];
    }
    off += tot;
  }
}

template <class Src> static void runScan(hipStream_t s, Src src, const uint32_t *nptr, uint32_t *partials) {
  static_assert(SCAN_BLOCKS <= 2 * TPB, "partials scan handles at most 2*TPB blocks");
  kScanReduce<Src><<<SCAN_BLOCKS, TPB, 0, s>>>(src, nptr, partials);
  kScanPartials<Src><<<1, TPB, 0, s>>>(src, nptr, partials, SCAN_BLOCKS);
  kScanDown<Src><<<SCAN_BLOCKS, TPB, 0, s>>>(src, nptr, partials);
}

// Candidate-count scan over bitmap words: cpre = exclusive prefix of cnt,
// ncand = base + total (clamped to the workspace capacity, with the
// overflow flag raised when it does not fit).
struct WordCountScan {
  const uint32_t *cnt;
  uint32_t *cpre;
  uint32_t *ncand;
  uint32_t nwords;
  int base;
  size_t ccap;
  DevScalars *sc;
  __device__ uint32_t size(const uint32_t *) const { return nwords; }
  __device__ uint32_t load(uint32_t i) const { return cnt[i]; }
  __device__ void store(uint32_t i, uint32_t p, uint32_t) const { cpre[i] = p; }
  __device__ void total(const uint32_t *, uint32_t t) const {
    cpre[nwords] = t;
    uint64_t nc = (uint64_t)t + (uint64_t)base;
    if (nc > ccap) {
      sc->overflow = 1;
      nc = ccap;
    }
    *ncand = (uint32_t)nc;
  }
};

// Keep-scan over candidates with scatter into the output band.
struct KeepScan {
  const double *out;
  const uint32_t *cpos;
  uint32_t *ks;
  double *val;
  uint32_t *pos;
  uint32_t *nOut;
  size_t cap;
  DevScalars *sc;
  __device__ uint32_t size(const uint32_t *nptr) const { return *nptr; }
  __device__ uint32_t load(uint32_t i) const { return keepFlag(out, i); }
  __device__ void store(uint32_t i, uint32_t p, uint32_t k) const {
    ks[i] = p;
    if (k && p < cap) {
      val[p] = out[i];
      pos[p] = cpos[i];
    }
  }
  __device__ void total(const uint32_t *nptr, uint32_t t) const {
    ks[*nptr] = t;
    if (t > cap) {
      sc->overflow = 1;
      t = (uint32_t)cap;
    }
    *nOut = t;
  }
};

// ------------------------------------------------------ candidate bits ----

template <int D>
__global__ void kCandStar(const uint64_t *bits, Frame f, GridInfo g, uint64_t *cbits, uint32_t *cnt, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const int64_t nw = f.nwords;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nw; w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t C = candStarWord<D>(bits, f, g, w);
    cbits[w] = C;
    cnt[w] = (uint32_t)__builtin_popcountll(C);
  }
}

// Union of two bitmaps (BooleanOperation visit set).
__global__ void kCandUnion(const uint64_t *a, const uint64_t *b, int64_t nw, uint64_t *cbits, uint32_t *cnt,
                           DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nw; w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t C = a[w] | b[w];
    cbits[w] = C;
    cnt[w] = (uint32_t)__builtin_popcountll(C);
  }
}

// Expand candidate bits into the position list (index base + cpre[w] + j).
__global__ void kCandList(const uint64_t *cbits, const uint32_t *cpre, int64_t nw, int base, uint32_t *cpos,
                          size_t ccap, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < nw; w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t C = cbits[w];
    if (w == 0 && base) cpos[0] = 0xFFFFFFFFu;
    if (!C) continue;
    size_t t = (size_t)base + cpre[w];
    if (t + (size_t)__builtin_popcountll(C) > ccap) {
      sc->overflow = 1;
      continue;
    }
    while (C) {
      int b = __builtin_ctzll(C);
      cpos[t++] = (uint32_t)((w << 6) | b);
      C &= C - 1;
    }
  }
}

// Bitmap words and prefix counts of the compacted output.
__global__ void kBuildIndex(const uint32_t *cpre, const uint32_t *ks, const uint32_t *cpos, int64_t nw, int base,
                            const uint32_t *ncand, uint64_t *bitsOut, uint32_t *preOut, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t nc = *ncand;
  for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w <= nw; w += (int64_t)gridDim.x * blockDim.x) {
    uint64_t word;
    uint32_t pre;
    buildIndexWord(cpre, ks, cpos, w, nw, base, nc, word, pre);
    bitsOut[w] = word;
    preOut[w] = pre;
  }
}

// ------------------------------------------------------------- rules ----

template <int D>
__global__ void kEvalStar(BandView in, const uint32_t *cpos, const uint32_t *ncand, OpParams op, double *out,
                          DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *ncand;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int32_t c[3];
    candCoords(in, i, cpos[i], c);
    double center, nb[2 * D];
    gatherStar<D>(in, c, center, nb);
    out[i] = (op.kind == OP_EXPAND) ? ruleExpand<D>(center, nb, op.a, op.b) : ruleRebuild<D>(center, nb, op.a);
  }
}

template <int D> __global__ void kEvalSelf(BandView in, const uint32_t *nptr, OpParams op, double *out, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double c = in.val[i];
    double r;
    if (op.kind == OP_REDUCE) {
      r = ruleReduce(c, op.a);
    } else if (op.kind == OP_UPDATE) {
      r = ruleUpdate(c, *op.dt, op.rates[i]);
    } else {
      int32_t cc[3];
      candCoords(in, i, in.pos[i], cc);
      double center, nb[2 * D];
      gatherStar<D>(in, cc, center, nb);
      r = rulePrune<D>(c, nb, op.kind == OP_PRUNE_RSZ);
    }
    out[i] = r;
  }
}

__global__ void kEvalBoolean(BandView A, BandView B, const uint32_t *cpos, const uint32_t *ncand, int kind, double wT,
                             double wS, double *out, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *ncand;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int32_t c[3];
    candCoords(A, i, cpos[i], c);
    double a = A.value(c[0], c[1], c[2]);
    double b = B.value(c[0], c[1], c[2]);
    out[i] = ruleBoolean(kind, a, b, wT, wS);
  }
}

// ------------------------------------------------------- host wrappers ----

static void compactInto(hipStream_t s, const Workspace &ws, const uint32_t *ncandPtr, const uint32_t *cpos,
                        const uint32_t *cpre, const Frame &f, BandBufs &out, DevScalars *sc) {
  KeepScan k;
  k.out = ws.out;
  k.cpos = cpos;
  k.ks = ws.ks;
  k.val = out.val;
  k.pos = out.pos;
  k.nOut = out.n;
  k.cap = out.cap;
  k.sc = sc;
  runScan(s, k, ncandPtr, ws.partials);
  kBuildIndex<<<gridFor((size_t)f.nwords + 1), TPB, 0, s>>>(cpre, ws.ks, cpos, f.nwords, f.base, ncandPtr, out.bits,
                                                            out.pre, sc);
}

static void candidatesFromBits(hipStream_t s, Workspace &ws, const Frame &f, DevScalars *sc) {
  WordCountScan wc;
  wc.cnt = ws.cnt;
  wc.cpre = ws.cpre;
  wc.ncand = ws.ncand;
  wc.nwords = (uint32_t)f.nwords;
  wc.base = f.base;
  wc.ccap = ws.ccap;
  wc.sc = sc;
  runScan(s, wc, nullptr, ws.partials);
  kCandList<<<gridFor((size_t)f.nwords), TPB, 0, s>>>(ws.cbits, ws.cpre, f.nwords, f.base, ws.cpos, ws.ccap, sc);
}

void launchStarOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc) {
  const int gw = gridFor((size_t)f.nwords);
  if (D == 2) kCandStar<2><<<gw, TPB, 0, s>>>(in.bits, f, g, ws.cbits, ws.cnt, sc);
  else kCandStar<3><<<gw, TPB, 0, s>>>(in.bits, f, g, ws.cbits, ws.cnt, sc);
  candidatesFromBits(s, ws, f, sc);
  BandView v = viewOf(in, f, g);
  const int ge = gridFor(ws.ccap, 8192);
  if (D == 2) kEvalStar<2><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, op, ws.out, sc);
  else kEvalStar<3><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, op, ws.out, sc);
  compactInto(s, ws, ws.ncand, ws.cpos, ws.cpre, f, out, sc);
}

void launchSelfOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc) {
  BandView v = viewOf(in, f, g);
  const int ge = gridFor(in.cap, 8192);
  if (D == 2) kEvalSelf<2><<<ge, TPB, 0, s>>>(v, in.n, op, ws.out, sc);
  else kEvalSelf<3><<<ge, TPB, 0, s>>>(v, in.n, op, ws.out, sc);
  compactInto(s, ws, in.n, in.pos, in.pre, f, out, sc);
}

void launchBoolean(hipStream_t s, int D, const BandBufs &A, const BandBufs &B, BandBufs &out, const Frame &f,
                   const GridInfo &g, Workspace &ws, int kind, double wTarget, double wSource, DevScalars *sc) {
  (void)D;
  kCandUnion<<<gridFor((size_t)f.nwords), TPB, 0, s>>>(A.bits, B.bits, f.nwords, ws.cbits, ws.cnt, sc);
  candidatesFromBits(s, ws, f, sc);
  BandView va = viewOf(A, f, g), vb = viewOf(B, f, g);
  const int ge = gridFor(ws.ccap, 8192);
  kEvalBoolean<<<ge, TPB, 0, s>>>(va, vb, ws.cpos, ws.ncand, kind, wTarget, wSource, ws.out, sc);
  compactInto(s, ws, ws.ncand, ws.cpos, ws.cpre, f, out, sc);
}

// ------------------------------------------------------------ rates ----

// Block max of non-negative doubles (as ordered bit patterns), then one
// atomicMax per block.
__device__ inline void blockMaxAtomic(unsigned long long v, unsigned long long *dst) {
  __shared__ unsigned long long sm[TPB / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    unsigned long long y = __shfl_down(v, o, 64);
    v = (y > v) ? y : v;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sm[0];
    for (int k = 1; k < TPB / 64; ++k) m = (sm[k] > m) ? sm[k] : m;
    if (m) atomicMax(dst, m);
  }
  __syncthreads();
}

__device__ inline unsigned long long blockSumU64(unsigned long long v) {
  __shared__ unsigned long long sm2[TPB / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm2[wid] = v;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int k = 0; k < TPB / 64; ++k) t += sm2[k];
  __syncthreads();
  return t;
}

// Hamiltonian + dissipation per active entry of the expanded band
// (Advect::integrateTime, lsAdvect.hpp:495-593, one level set): writes the
// stage velocity and reduces max |H - D| (and the LLF/LLLF alphas).
template <int D, int SCHEME>
__global__ void kRates(BandView b, const uint32_t *nptr, RateParams rp, double *rates, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  const Vel vel = velOf(rp);
  double lmax = 0.;
  double la[3] = {0, 0, 0};
  double ga[3] = {0, 0, 0};
  unsigned long long act = 0;
  if (SCHEME == S_LF)
    for (int i = 0; i < 3; ++i) ga[i] = __longlong_as_double((long long)sc->alphaBits[i]);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double phi0 = b.val[i];
    if (isUndef(phi0) || fabs(phi0) > 0.5) continue;
    ++act;
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    double velocity;
    rates[i] = ratePoint<D, SCHEME>(b, c, phi0, rp, vel, ga, la, velocity);
    if (velocity != 0.) {
      double a = fabs(velocity);
      if (lmax < a) lmax = a;
    }
  }
  blockMaxAtomic((unsigned long long)__double_as_longlong(lmax), &sc->maxAbsBits);
  if (SCHEME == S_LLLF || SCHEME == S_LLF)
    for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(la[d]), &sc->alphaBits[d]);
  unsigned long long t = blockSumU64(act);
  if (threadIdx.x == 0 && t) atomicAdd(&sc->activeCount, t);
}

__global__ void kTimeStep(DevScalars *sc, RateParams rp, int D, double dx, const double *maxTimeStepPtr, int useAlphas,
                          double *dtOut, const uint32_t *nb) {
  LSG_ABORT_IF_OVERFLOW(sc);
  double al[3];
  for (int i = 0; i < 3; ++i) al[i] = __longlong_as_double((long long)sc->alphaBits[i]);
  const double m = __longlong_as_double((long long)sc->maxAbsBits);
  *dtOut = stageTimeStep(*maxTimeStepPtr, rp.cfl, m, D, al, useAlphas != 0, dx);
  sc->totalActive += sc->activeCount;
  // algorithmic bytes of the rate kernel: every band entry read once (the
  // stencil support lies in the band), one rate written per active point
  sc->rateBytes += 8.0 * (double)(*nb) + 8.0 * (double)sc->activeCount;
}

__global__ void kResetRates(DevScalars *sc, int keepAlphas) {
  sc->maxAbsBits = 0;
  if (!keepAlphas) sc->alphaBits[0] = sc->alphaBits[1] = sc->alphaBits[2] = 0;
  sc->activeCount = 0;
}

void launchRates(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const RateParams &rp,
                 double *rates, DevScalars *sc, const double *maxTimeStep, double *dtOut) {
  const int scheme = schemeIdOf(rp.spatial);
  kResetRates<<<1, 1, 0, s>>>(sc, scheme == S_LF ? 1 : 0);
  BandView v = viewOf(b, f, g);
  const int ge = gridFor(b.cap, 8192);
#define LSG_RATES(DD, SS) kRates<DD, SS><<<ge, TPB, 0, s>>>(v, b.n, rp, rates, sc)
  if (D == 2) {
    switch (scheme) {
    case S_EO: LSG_RATES(2, S_EO); break;
    case S_WENO5: LSG_RATES(2, S_WENO5); break;
    case S_LF: LSG_RATES(2, S_LF); break;
    case S_LLLF: LSG_RATES(2, S_LLLF); break;
    case S_LLF: LSG_RATES(2, S_LLF); break;
    default: LSG_RATES(2, S_WENO3); break;
    }
  } else {
    switch (scheme) {
    case S_EO: LSG_RATES(3, S_EO); break;
    case S_WENO5: LSG_RATES(3, S_WENO5); break;
    case S_LF: LSG_RATES(3, S_LF); break;
    case S_LLLF: LSG_RATES(3, S_LLLF); break;
    case S_LLF: LSG_RATES(3, S_LLF); break;
    default: LSG_RATES(3, S_WENO3); break;
    }
  }
#undef LSG_RATES
  const int useAlphas = (scheme == S_LF || scheme == S_LLLF || scheme == S_LLF) ? 1 : 0;
  kTimeStep<<<1, 1, 0, s>>>(sc, rp, D, g.dx, maxTimeStep, useAlphas, dtOut, b.n);
}

// Advect::findGlobalAlphas (lsAdvect.hpp:90-195), one level set.
template <int D> __global__ void kGlobalAlphas(BandView b, const uint32_t *nptr, RateParams rp, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  const Vel vel = velOf(rp);
  double la[3] = {0, 0, 0};
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double v0 = b.val[i];
    if (isUndef(v0) || fabs(v0) > 0.5) continue;
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    globalAlphaPoint<D>(b, c, vel, la);
  }
  for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(la[d]), &sc->alphaBits[d]);
}

__global__ void kResetAlphas(DevScalars *sc) { sc->alphaBits[0] = sc->alphaBits[1] = sc->alphaBits[2] = 0; }

void launchGlobalAlphas(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g,
                        const RateParams &rp, DevScalars *sc) {
  kResetAlphas<<<1, 1, 0, s>>>(sc);
  BandView v = viewOf(b, f, g);
  const int ge = gridFor(b.cap, 8192);
  if (D == 2) kGlobalAlphas<2><<<ge, TPB, 0, s>>>(v, b.n, rp, sc);
  else kGlobalAlphas<3><<<ge, TPB, 0, s>>>(v, b.n, rp, sc);
}

// Returned time of the step (lsAdvectIntegrationSchemes.hpp:58,107,167).
__global__ void kStepTime(DevScalars *sc, int temporal) {
  if (temporal == 1) sc->stepTime = 0.5 * sc->dt[0] + 0.5 * sc->dt[1];
  else if (temporal == 2) sc->stepTime = (sc->dt[0] + sc->dt[1] + 4.0 * sc->dt[2]) / 6.0;
  else sc->stepTime = sc->dt[0];
}
void launchStepTime(hipStream_t s, DevScalars *sc, int temporal) { kStepTime<<<1, 1, 0, s>>>(sc, temporal); }

// ---------------------------------------------------------------- misc ----

__global__ void kCopyBand(const double *sv, const uint32_t *sp, const uint32_t *sn, double *dv, uint32_t *dp,
                          uint32_t *dn, size_t cap, const uint64_t *sb, const uint32_t *spre, uint64_t *db,
                          uint32_t *dpre, int64_t nw) {
  uint32_t n = *sn;
  if (n > cap) n = (uint32_t)cap;
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < n; i += stride) {
    dv[i] = sv[i];
    dp[i] = sp[i];
  }
  for (size_t w = tid; w <= (size_t)nw; w += stride) {
    db[w] = sb[w];
    dpre[w] = spre[w];
  }
  if (tid == 0) *dn = n;
}
void launchCopyBand(hipStream_t s, const BandBufs &src, BandBufs &dst, const Frame &f) {
  size_t work = src.cap > (size_t)f.nwords ? src.cap : (size_t)f.nwords;
  kCopyBand<<<gridFor(work, 8192), TPB, 0, s>>>(src.val, src.pos, src.n, dst.val, dst.pos, dst.n, dst.cap, src.bits,
                                                src.pre, dst.bits, dst.pre, f.nwords);
}

__global__ void kBandStats(const double *val, const uint32_t *pos, const uint32_t *nptr, Frame f, uint32_t *count,
                           int32_t *bbox) {
  const uint32_t n = *nptr;
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  unsigned long long c = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!isUndef(val[i])) ++c;
    if (i == 0 && f.base) continue;
    int32_t r[3];
    f.decode(pos[i], r[0], r[1], r[2]);
    for (int d = 0; d < 3; ++d) {
      mn[d] = r[d] < mn[d] ? r[d] : mn[d];
      mx[d] = r[d] > mx[d] ? r[d] : mx[d];
    }
  }
  unsigned long long t = blockSumU64(c);
  if (threadIdx.x == 0 && t) atomicAdd(count, (uint32_t)t);
  for (int d = 0; d < 3; ++d) {
    if (mn[d] != INT32_MAX) atomicMin(&bbox[d], mn[d]);
    if (mx[d] != INT32_MIN) atomicMax(&bbox[3 + d], mx[d]);
  }
}
__global__ void kStatsInit(uint32_t *count, int32_t *bbox) {
  *count = 0;
  for (int d = 0; d < 3; ++d) {
    bbox[d] = INT32_MAX;
    bbox[3 + d] = INT32_MIN;
  }
}
void launchBandStats(hipStream_t s, const BandBufs &b, const Frame &f, uint32_t *outCount, int32_t *bbox) {
  kStatsInit<<<1, 1, 0, s>>>(outCount, bbox);
  kBandStats<<<gridFor(b.cap, 2048), TPB, 0, s>>>(b.val, b.pos, b.n, f, outCount, bbox);
}

__global__ void kFill32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
void launchFill32(hipStream_t s, uint32_t *p, uint32_t v, size_t n) { kFill32<<<gridFor(n), TPB, 0, s>>>(p, v, n); }

// Direct evaluation of FiniteDifferences::weno3/weno5 on n stencils.
__global__ void kWeno(int order, uint32_t n, const double *x, double dx, int plus, double *out) {
  const int w = (order == 3) ? 5 : 7;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double st[7];
    for (int k = 0; k < w; ++k) st[k] = x[(size_t)i * w + k];
    out[i] = (order == 3) ? weno3(st, dx, plus != 0) : weno5(st, dx, plus != 0);
  }
}
void launchWeno(hipStream_t s, int order, uint32_t n, const double *x, double dx, int plus, double *out) {
  kWeno<<<gridFor(n), TPB, 0, s>>>(order, n, x, dx, plus, out);
}

}  // namespace lsg
