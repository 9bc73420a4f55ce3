// HIP kernels of the level-set advection path, written for CDNA4 (gfx950,
// wave64).  Every kernel reads its problem size from device memory and runs
// a grid-stride loop over a fixed grid, so a whole time step is enqueued
// without host round trips.  The per-item logic lives in lsg_items.hpp.
//
// Once any kernel of a step flags a capacity overflow, every later kernel of
// the step returns immediately (the host then retries the step with larger
// buffers), so no kernel ever indexes past its buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>

#include "lsg_device.hpp"
#include "lsg_items.hpp"
#include "lsg_kernels.hpp"

namespace lsg {

static constexpr int TPB = 256;
// Grid of the scan kernels (blocks; at most SCAN_BLOCKS_MAX), chosen at run
// time (batch autotuning).  Default: SCAN_BLOCKS, or LSG_SCAN_BLOCKS from the
// environment (grid sweeps).
static int clampBlocks(int b) { return b < 8 ? 8 : (b > SCAN_BLOCKS_MAX ? SCAN_BLOCKS_MAX : b); }
int defaultScanBlocks() {
  static const int v = [] {
    const char *e = getenv("LSG_SCAN_BLOCKS");
    return e ? clampBlocks(atoi(e)) : SCAN_BLOCKS;
  }();
  return v;
}
static int g_scanBlocks = defaultScanBlocks();
// Grid of the rate kernel (4 workgroups per CU); LSG_RATE_BLOCKS from the
// environment overrides it (grid sweeps).
static int rateBlocks() {
  static const int v = [] {
    const char *e = getenv("LSG_RATE_BLOCKS");
    return e ? clampBlocks(atoi(e)) : 1024;
  }();
  return v;
}
void setScanBlocks(int b) { g_scanBlocks = clampBlocks(b); }
int scanBlocks() { return g_scanBlocks; }
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
  v.rb = (uint32_t)f.base;
  return v;
}

#define LSG_ABORT_IF_OVERFLOW(sc) \
  if ((sc)->overflow) return

// ------------------------------------------------------------- scans ----
//
// Two-kernel scans with device-side sizes.  Phase 1: every block reduces its
// contiguous chunk (and may produce the values, e.g. evaluate a rule) into
// partials[b].  Phase 2: every block sums the partials before it (all
// threads cooperate), rescans its chunk from that offset and scatters.  No
// block ever waits on another block.

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

__device__ inline uint32_t blockSum(uint32_t v) {
  uint32_t t;
  blockExclusiveScan(v, t);
  return t;
}

// Logical block index: workgroups are dispatched round-robin over the 8 XCDs
// (each with its own L2), so consecutive logical blocks are mapped to the
// same XCD; a contiguous 1/8 of the data then lives in one L2 and stencil
// neighbours (rows above and below) are found there.
static constexpr uint32_t NUM_XCD = 8;
__device__ inline uint32_t logicalBlock() {
  const uint32_t nb = gridDim.x, b = blockIdx.x;
  const uint32_t xcd = b % NUM_XCD, k = b / NUM_XCD;
  const uint32_t per = nb / NUM_XCD, rem = nb % NUM_XCD;
  return xcd * per + (xcd < rem ? xcd : rem) + k;
}

// Contiguous chunk [lo, hi) of the (logical) block for a device-side size n;
// chunks are multiples of `align` items (the packed keep flags of the
// compaction are written one 64-bit word per wave).
__device__ inline uint32_t chunkSize(uint32_t n, uint32_t align) {
  const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
  return (per + align - 1) / align * align;
}
__device__ inline void blockChunk(uint32_t n, uint32_t &lo, uint32_t &hi, uint32_t align = 64) {
  const uint32_t per = chunkSize(n, align);
  lo = logicalBlock() * per;
  hi = lo + per;
  if (lo > n) lo = n;
  if (hi > n) hi = n;
}

// Sum of partials[0 .. logical block) computed by the whole block.  The
// partials (SCAN_BLOCKS_MAX + 1 counts, 256-byte aligned) are read four at a
// time, every thread's loads issued before the first is used (a loop that
// adds each count as it arrives waits once per load).
__device__ inline uint32_t blockOffset(const uint32_t *partials) {
  static_assert(SCAN_BLOCKS_MAX <= 4 * 4 * TPB, "four uint4 loads per thread cover the partials");
  const uint32_t me = logicalBlock();
  const uint4 *p4 = reinterpret_cast<const uint4 *>(partials);
  uint4 v[4];
  LSG_UNROLL
  for (int r = 0; r < 4; ++r) {
    const uint32_t q = threadIdx.x + (uint32_t)r * TPB;
    v[r] = (4u * q < me) ? p4[q] : make_uint4(0u, 0u, 0u, 0u);
  }
  uint32_t s = 0;
  LSG_UNROLL
  for (int r = 0; r < 4; ++r) {
    const uint32_t k = 4u * (threadIdx.x + (uint32_t)r * TPB);
    s += (k < me ? v[r].x : 0u) + (k + 1 < me ? v[r].y : 0u) + (k + 2 < me ? v[r].z : 0u) + (k + 3 < me ? v[r].w : 0u);
  }
  return blockSum(s);
}

// Value of the previous item in thread order within a tile: lane-1 of the
// wave, the last lane of the previous wave, or `carry` for thread 0.  With
// tileLast, every thread also receives the tile's last item.
__device__ inline double prevInTile(double last, double carry, double *tileLast = nullptr) {
  __shared__ double wl[TPB / 64];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  double up = __shfl_up(last, 1, 64);
  if (lane == 63) wl[wid] = last;
  __syncthreads();
  double p = (lane > 0) ? up : (wid > 0 ? wl[wid - 1] : carry);
  if (tileLast) *tileLast = wl[TPB / 64 - 1];
  __syncthreads();
  return p;
}

// ------------------------------------------------------ candidate bits ----

// Phase 1 of candidate generation: candidate bitmap word per word (star
// dilation or union) and per-block candidate counts.
template <int D, int MODE, bool SEG>  // MODE 0: star dilation of A; 1: union of A and B
__global__ __launch_bounds__(256) void kCandCount(const uint64_t *A, const uint64_t *B, Frame f, GridInfo g, uint32_t nw,
                                                  uint64_t *cbits, uint32_t *partials, DevScalars *sc, SegArgs sa) {
  if (sc->overflow) return;
  uint32_t lo, hi;
  blockChunk(nw, lo, hi);
  uint32_t acc = 0;
  for (uint32_t w = lo + threadIdx.x; w < hi; w += TPB) {
    uint64_t C;
    if (MODE == 1) {
      C = A[w] | B[w];
    } else if (!SEG) {
      C = candStarWord<D>(A, f, g, (int64_t)w);
    } else {
      const SegInfo &S = sa.segs[sa.wseg[w]];
      C = candStarWord<D>(A + S.wordOff, S.f, S.g, (int64_t)(w - S.wordOff));
    }
    cbits[w] = C;
    acc += (uint32_t)__builtin_popcountll(C);
  }
  uint32_t t = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = t;
}

// Phase 2: word prefix counts and the candidate position list (index
// base + cpre[w] + j); the last block stores the totals.  The positions of a
// word are written by the 64 lanes of a wave at once (lane = bit), so the
// stores of a full word are one contiguous 256-byte segment.  In a batch every level set's
// candidates are preceded by its grid-minimum candidate (when its frame
// does not hold the grid minimum), so the candidate index of a word is
// shifted by the grid-minimum entries up to its level set, and each
// candidate records its level set in cseg.
template <bool SEG>
__global__ __launch_bounds__(256) void kCandList(const uint64_t *cbits, uint32_t nw, int base, const uint32_t *partials,
                                                 uint32_t *cpre, uint32_t *cpos, size_t ccap, uint32_t *ncand,
                                                 DevScalars *sc, SegArgs sa, uint16_t *cseg) {
  uint32_t lo, hi;
  blockChunk(nw, lo, hi);
  // a tile's words are loaded one tile ahead; the first tile's loads are
  // issued before the block offset is summed
  uint64_t cNext = (lo + threadIdx.x < hi) ? cbits[lo + threadIdx.x] : 0ull;
  uint32_t off = blockOffset(partials);
  if (sc->overflow) return;  // uniform: after the block-wide sum
  if (!SEG && logicalBlock() == 0 && threadIdx.x == 0 && base) cpos[0] = POS_NONE;
  const int lane = threadIdx.x & 63;
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t w = t + threadIdx.x;
    const uint64_t c = cNext;
    if (t + TPB < hi) cNext = (w + TPB < hi) ? cbits[w + TPB] : 0ull;
    const uint32_t pc = (uint32_t)__builtin_popcountll(c);
    uint32_t tot;
    const uint32_t pl = blockExclusiveScan(pc, tot);
    if (w < hi) cpre[w] = off + pl;
    uint32_t sh = 0, lw = 0, k = 0;
    if (SEG && w < hi) {
      k = sa.wseg[w];
      const SegInfo &S = sa.segs[k];
      sh = S.baseIncl;
      lw = w - S.wordOff;  // positions are level-set local
      if (w == S.wordOff && S.f.base) {  // the level set's grid-minimum candidate precedes its first word
        const size_t q = (size_t)off + pl + sh - 1;
        if (q < ccap) {
          cpos[q] = POS_NONE;
          cseg[q] = (uint16_t)k;
        }
      }
    }
    // a full tile gives each wave its own 64 words; a partial one (a chunk
    // shorter than the tile: 65 words per block at the trench256 size) is
    // dealt to the waves through LDS as contiguous ranges of ceil(valid / 4)
    // words, so no wave is left with all of them.  Each wave then writes the
    // positions of its non-empty words one word per iteration (read from the
    // owning lanes) with lane = bit, so consecutive lanes write consecutive
    // slots.  Slot arithmetic is 32-bit: candidates are distinct frame bits
    // (< 2^32) and the candidate capacity is below 2^30, so q < ccap also
    // bounds the byte offset.
    __shared__ uint64_t tileWord[TPB];
    __shared__ uint32_t tilePre[TPB];
    const uint32_t valid = (hi - t < (uint32_t)TPB) ? hi - t : (uint32_t)TPB;  // uniform
    uint64_t cs = c;
    uint32_t ps = pl, slot = threadIdx.x;
    if (valid < (uint32_t)TPB) {
      tileWord[threadIdx.x] = (w < hi) ? c : 0ull;
      tilePre[threadIdx.x] = pl;
      __syncthreads();
      const uint32_t per = (valid + TPB / 64 - 1) / (TPB / 64);
      slot = (uint32_t)(threadIdx.x >> 6) * per + (uint32_t)lane;
      const bool mine = (uint32_t)lane < per && slot < valid;
      cs = mine ? tileWord[slot] : 0ull;
      ps = mine ? tilePre[slot] : 0u;
      __syncthreads();  // the next tile rewrites tileWord / tilePre
    }
    uint32_t shs = 0, lws = t + slot, kss = 0;
    if (SEG && cs) {
      kss = sa.wseg[t + slot];
      const SegInfo &S = sa.segs[kss];
      shs = S.baseIncl;
      lws = t + slot - S.wordOff;  // positions are level-set local
    }
    const uint32_t q0 = (SEG ? 0u : (uint32_t)base) + off;
    const uint32_t cap32 = (uint32_t)ccap;
    const uint32_t lsh = (uint32_t)lane & 31u;
    const uint32_t belowLo = lane < 32 ? ((1u << lsh) - 1u) : 0xFFFFFFFFu;
    const uint32_t belowHi = lane < 32 ? 0u : ((1u << lsh) - 1u);
    uint64_t todo = (uint64_t)__ballot(cs != 0ull);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1ull;
      const uint32_t cwLo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)cs, j);
      const uint32_t cwHi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(cs >> 32), j);
      const uint32_t pj = (uint32_t)__builtin_amdgcn_readlane((int)ps, j);
      const uint32_t word = (uint32_t)__builtin_amdgcn_readlane((int)lws, j);
      const uint32_t shj = SEG ? (uint32_t)__builtin_amdgcn_readlane((int)shs, j) : 0u;
      if ((((lane < 32) ? cwLo : cwHi) >> lsh) & 1u) {
        const uint32_t q = q0 + pj + (uint32_t)__builtin_popcount(cwLo & belowLo) +
                           (uint32_t)__builtin_popcount(cwHi & belowHi) + shj;
        if (q < cap32) {
          *(uint32_t *)((char *)cpos + (q << 2)) = (word << 6) | (uint32_t)lane;
          if (SEG) *(uint16_t *)((char *)cseg + (q << 1)) = (uint16_t)__builtin_amdgcn_readlane((int)kss, j);
        }
      }
    }
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) {
    cpre[nw] = off;
    uint64_t nc = (uint64_t)off + (SEG ? (uint64_t)sa.totalBase : (uint64_t)base);
    if (nc > ccap) {
      sc->overflow = 1;
      nc = ccap;
    }
    *ncand = (uint32_t)nc;
  }
}

// ------------------------------------------------- evaluators of rules ----

template <int D, bool SEG> struct EvalStar {
  static constexpr int kCount = SEG ? 0 : 1;  // single band: algorithmic bytes counted in starBytes
  BandView in;
  SegArgs sa;
  const uint32_t *cpos;
  OpParams op;
  // entries of the input band (in-frame ones plus the grid-minimum entry)
  __device__ uint32_t inputEntries() const { return in.pre[in.f.nwords] + in.rb; }
  __device__ double operator()(uint32_t i) const {
    const BandView v = SEG ? segView(in, sa.segs[sa.iseg[i]]) : in;
    int32_t c[3];
    candCoords(v, i, cpos[i], c);
    double center, nb[2 * D];
    gatherStar<D>(v, c, center, nb);
    return (op.kind == OP_EXPAND) ? ruleExpand<D>(center, nb, op.a, op.b) : ruleRebuild<D>(center, nb, op.a);
  }
};

// Two Expand cycles (OP_EXPAND2) at every candidate of the twice-dilated
// candidate bitmap (one level set).
template <int D> struct EvalStar2 {
  static constexpr int kCount = 2;  // algorithmic bytes counted in star2Bytes
  BandView in;
  const uint32_t *cpos;
  OpParams op;
  __device__ uint32_t inputEntries() const { return in.pre[in.f.nwords] + in.rb; }
  // inlined into the kernel: as a call it passed the evaluator through
  // scratch memory
  __device__ __attribute__((always_inline)) double operator()(uint32_t i) const {
    int32_t c[3];
    candCoords(in, i, cpos[i], c);
    return expandTwice<D>(in, c, op.a, op.a2, op.b);
  }
};

// updateLevelSet and rebuildLS (OP_UPDATE_REBUILD) at every candidate of
// the star dilation of the band the rates belong to (one level set).
template <int D> struct EvalUpdRebuild {
  static constexpr int kCount = 0;
  BandView in;
  const uint32_t *cpos;
  OpParams op;
  __device__ uint32_t inputEntries() const { return 0; }
  __device__ __attribute__((always_inline)) double operator()(uint32_t i) const {
    int32_t c[3];
    candCoords(in, i, cpos[i], c);
    return updateRebuild<D>(in, c, op.rates, *op.dt, op.a);
  }
};

template <int D, bool SEG> struct EvalSelf {
  static constexpr int kCount = 0;
  __device__ uint32_t inputEntries() const { return 0; }
  BandView in;
  SegArgs sa;
  OpParams op;
  __device__ double operator()(uint32_t i) const {
    double c = in.val[i];
    const uint32_t k = SEG ? sa.iseg[i] : 0u;
    if (op.kind == OP_REDUCE) return ruleReduce(c, op.a);
    if (op.kind == OP_UPDATE) return ruleUpdate(c, op.dt[k], op.rates[i]);
    if (op.kind == OP_UPDATE_MM) return ruleUpdateMM(c, op.dt[k], op.rates + (size_t)i * 2 * op.nmat);
    const BandView v = SEG ? segView(in, sa.segs[k]) : in;
    int32_t cc[3];
    candCoords(v, i, v.pos[i], cc);
    double center, nb[2 * D];
    gatherStar<D>(v, cc, center, nb);
    return rulePrune<D>(c, nb, op.kind == OP_PRUNE_RSZ);
  }
};

template <bool SEG> struct EvalBool {
  static constexpr int kCount = 0;
  __device__ uint32_t inputEntries() const { return 0; }
  BandView A, B;
  SegArgs sa;
  const uint32_t *cpos;
  int kind;
  double wT, wS;
  uint32_t *movedDown;  // combine only: set where the result exceeds b (both defined)
  __device__ double operator()(uint32_t i) const {
    BandView va = A, vb = B;
    if (SEG) {
      const SegInfo &S = sa.segs[sa.iseg[i]];
      va = segView(A, S);
      vb = segView(B, S);
    }
    int32_t c[3];
    candCoords(va, i, cpos[i], c);
    BandView::Loc l[2] = {va.locate(c[0], c[1], c[2]), vb.locate(c[0], c[1], c[2])};
    const uint64_t wa = va.bits[l[0].w], wb = vb.bits[l[1].w];
    const uint32_t pa = va.pre[l[0].w], pb = vb.pre[l[1].w];
    const double a = va.val[va.entryOf(l[0], wa, pa)];
    const double b = vb.val[vb.entryOf(l[1], wb, pb)];
    const double r = ruleBoolean(kind, a, b, wT, wS);
    if (movedDown && kind == B_COMBINE && !isUndef(a) && !isUndef(b) && r > b) *movedDown = 1u;
    return r;
  }
};

// Block-wide exclusive scan of one 0/1 flag per thread via wave ballots.
__device__ inline uint32_t blockBallotScan(bool k, uint32_t &total, uint64_t &waveMask) {
  __shared__ uint32_t wc[TPB / 64];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint64_t m = (uint64_t)__ballot(k ? 1 : 0);
  const uint32_t lp = (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wc[wid] = (uint32_t)__builtin_popcountll(m);
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < TPB / 64; ++q) {
    uint32_t c = wc[q];
    if (q < wid) woff += c;
    tot += c;
  }
  __syncthreads();
  total = tot;
  waveMask = m;
  return woff + lp;
}

// Keep flag of a candidate from its value and the previous candidate's: the
// first candidate of a level set, a defined value, or the start of a run
// whose value differs from the previous candidate's.
__device__ inline uint32_t keepOf(bool first, double v, double prev) {
  if (first) return 1;
  if (!isUndef(v)) return 1;
  return v != prev ? 1u : 0u;
}
template <bool SEG> __device__ inline bool firstOfSet(uint32_t i, const uint16_t *iseg) {
  if (i == 0) return true;
  return SEG ? (iseg[i] != iseg[i - 1]) : false;
}

// Phase 1 of compaction: evaluate the rule at every candidate (out[i]) and
// count the kept candidates per block.  A block's first keep flag needs the
// value of the candidate before its chunk: the block evaluates its chunk
// from that candidate on (evaluated again, not stored), so no thread waits
// on a separate evaluation of it.
template <class Ev, bool SEG>
__global__ __launch_bounds__(256) void kEvalKeep(Ev ev, const uint32_t *nptr, double *out, uint32_t *partials,
                                                 DevScalars *sc, const uint16_t *iseg) {
  const uint32_t n = *nptr;
  if (sc->overflow) return;
  uint32_t lo, hi;
  blockChunk(n, lo, hi);
  const uint32_t from = (lo > 0 && lo < hi) ? lo - 1 : lo;
  double carry = 0.0;  // last value of the previous tile
  uint32_t acc = 0;
  // one tile at a time: the larger evaluators hold many loads in flight
#pragma unroll 1
  for (uint32_t t = from; t < hi; t += TPB) {
    const uint32_t i = t + threadIdx.x;
    const bool ok = i >= lo && i < hi;
    const double v = ev(i < hi ? i : hi - 1);  // past the chunk: evaluated, not stored
    if (ok) out[i] = v;
    const double prev = prevInTile(v, carry, &carry);
    if (ok) acc += keepOf(firstOfSet<SEG>(i, iseg), v, prev);
  }
  const uint32_t tot = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = tot;
  // algorithmic bytes of a star operation's evaluation: every input entry
  // read once (the star supports lie in the band), one value written per
  // candidate
  if constexpr (Ev::kCount == 1)
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->starBytes += 8.0 * ((double)ev.inputEntries() + (double)n);
  if constexpr (Ev::kCount == 2)
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->star2Bytes += 8.0 * ((double)ev.inputEntries() + (double)n);
}

// Phase 2: rescan the keep flags from the block offset, scatter the kept
// candidates into the output band, store the running index ks[] and the
// keep flags packed one bit per candidate; the last block stores totals.
template <bool SEG>
__global__ __launch_bounds__(256) void kKeepScatter(const double *out, const uint32_t *cpos, const uint32_t *nptr,
                                                    const uint32_t *partials, uint32_t *ks, uint64_t *keepBits,
                                                    double *val, uint32_t *pos, uint32_t *nOut, size_t cap,
                                                    DevScalars *sc, const uint16_t *iseg, uint16_t *oseg) {
  const uint32_t n = *nptr;
  if (sc->overflow) return;
  uint32_t lo, hi;
  blockChunk(n, lo, hi);
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  // a tile's values are loaded one tile ahead; the first tile's loads are
  // issued before the block offset is summed
  double vNext = 0.0, pNext = 0.0;
  {
    const uint32_t i = lo + threadIdx.x;
    if (i < hi) {
      vNext = out[i];
      pNext = (i > 0) ? out[i - 1] : 0.0;
    }
  }
  uint32_t off = blockOffset(partials);
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t i = t + threadIdx.x;
    const bool valid = i < hi;
    const double v = vNext, prev = pNext;
    if (t + TPB < hi) {
      const uint32_t j = i + TPB;
      if (j < hi) {
        vNext = out[j];
        pNext = out[j - 1];
      }
    }
    uint32_t k = 0;
    if (valid) k = keepOf(firstOfSet<SEG>(i, iseg), v, prev);
    uint32_t tot;
    uint64_t wm;
    const uint32_t p = off + blockBallotScan(k != 0, tot, wm);
    if (valid) {
      ks[i] = p;
      if (k && p < cap) {
        val[p] = v;
        pos[p] = cpos[i];
        if (SEG) oseg[p] = iseg[i];
      }
    }
    const uint32_t g = t + (uint32_t)wid * 64u;  // 64-aligned: chunks and tiles are multiples of 256
    if (lane == 0 && g < hi) keepBits[g >> 6] = wm;
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) {
    ks[n] = off;
    uint32_t tt = off;
    if (tt > cap) {
      sc->overflow = 1;
      tt = (uint32_t)cap;
    }
    *nOut = tt;
  }
}

// Bitmap words and prefix counts of the compacted output: the kept flags of
// the candidates of word w (a contiguous run of the packed keep bits) are
// deposited onto the set bits of the candidate word.  The candidate index of
// word w is cpre[w] plus the grid-minimum candidates before it (base, or in
// a batch those of the level sets up to w's).
template <bool SEG>
__global__ void kBuildIndex(const uint64_t *cwords, const uint32_t *cpre, const uint32_t *ks, const uint64_t *keepBits,
                            uint32_t nw, int base, const uint32_t *ncand, uint64_t *bitsOut, uint32_t *preOut,
                            DevScalars *sc, SegArgs sa) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t nc = *ncand;
  uint32_t lo, hi;
  blockChunk(nw + 1, lo, hi);
  for (uint32_t w = lo + threadIdx.x; w < hi; w += TPB) {
    uint32_t sh = (uint32_t)base;
    if (SEG) sh = (w < nw) ? sa.segs[sa.wseg[w]].baseIncl : sa.totalBase;
    uint32_t c0 = sh + cpre[w];
    if (c0 > nc) c0 = nc;
    preOut[w] = ks[c0] - sh;
    if (w == nw) {
      bitsOut[w] = 0;
      continue;
    }
    const uint64_t src = cwords[w];
    uint64_t word = 0;
    if (src) {
      const uint32_t len = (uint32_t)__builtin_popcountll(src);
      const uint32_t q = c0 >> 6, s = c0 & 63;
      uint64_t m = keepBits[q] >> s;
      if (s != 0 && s + len > 64) m |= keepBits[q + 1] << (64 - s);
      if (len < 64) m &= (1ull << len) - 1ull;
      word = (src == ~0ull) ? m : depositBits(m, src);  // keep flags onto the candidate bits
    }
    bitsOut[w] = word;
  }
}

// ------------------------------------------- single-pass band operation ----
//
// One launch per band operation of a single level set: workgroup t takes
// tile t (BAND_TILE_WORDS bitmap words), builds its candidate words,
// evaluates the rule at every candidate, compacts the kept results in LDS,
// obtains the number of entries of all earlier tiles by decoupled look-back
// over per-tile status words (Merrill & Garland, "Single-pass Parallel Prefix
// Scan with Decoupled Look-back", 2016), and writes the tile's entries,
// bitmap words and prefix counts.  A tile waits only on tiles of lower
// workgroup index; workgroups are dispatched to each XCD in increasing index
// order, so the smallest unfinished tile is always running and never waits.
// Each wait is also bounded: when it runs out the kernel flags lookbackFail
// and the host turns the single-pass kernels off and repeats the work with
// the multi-kernel pipeline below.  No counter is shared between workgroups:
// a global claim counter costs ~25 ns per atomic when all XCDs hit one
// address (measured: 185 us for 7200 tiles).
//
// Keep flags: an entry is kept when it is defined or starts a run whose value
// differs from the previous candidate's (keepOf).  Slot 0 of a tile holds the
// rule evaluated at the grid point before the tile's first candidate: that is
// the previous candidate's result (the At* evaluators' contract), so no tile
// reads another tile's values.
enum CandMode : int { CAND_STAR = 0, CAND_UNION = 1, CAND_SELF = 2 };

static constexpr int TWS = BAND_TILE_WORDS;  // tile width of the batch kernel
static constexpr uint32_t LOOKBACK_SPIN_LIMIT = 1u << 15;  // about 30 ms of polling
static constexpr int LOOKBACK_PER_LANE = 8;  // a look-back window spans 64 x 8 tiles

__device__ inline uint32_t launchTag(uint32_t epoch) { return (epoch % 0x3FFFFFFFu) + 1u; }  // never 0
__device__ inline uint64_t tileWord(uint32_t flag, uint32_t tag, uint32_t v) {
  return ((uint64_t)flag << 62) | ((uint64_t)tag << 32) | (uint64_t)v;
}
constexpr uint32_t TILE_AGG = 1, TILE_INCL = 2;

// Exclusive prefix of the counts of the tiles before tile t (one wave, every
// lane active); publishes tile t's aggregate, then its inclusive count.  Lane
// l reads the tiles look - 1 - (l + 64 q), q < LOOKBACK_PER_LANE, so one
// round of loads covers 512 predecessors; the prefix is complete when every
// tile up to the nearest inclusive count is ready.
// Look-back tuning: the first poll reads 64 predecessors; while none of
// them holds an inclusive count (and all are ready) the window grows to
// 64 x lbq tiles; between polls a wave sleeps lbSleep x 128 clocks.
struct Lookback {
  uint32_t q;      // 1..LOOKBACK_PER_LANE
  uint32_t sleep;  // >= 1
};
__device__ uint32_t lookBack(uint64_t *stat, uint32_t t, uint32_t agg, uint32_t tag, DevScalars *sc, Lookback lb) {
  const int lane = threadIdx.x & 63;
  if (lane == 0)
    __hip_atomic_store(&stat[t], tileWord(t == 0 ? TILE_INCL : TILE_AGG, tag, agg), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  if (t == 0) return 0;
  constexpr int Q = LOOKBACK_PER_LANE;
  constexpr uint32_t NONE = 0xFFFFFFFFu;
  uint32_t excl = 0;
  int64_t look = t;  // the window holds tiles look - 1 - j, j < 64 qw
  uint32_t qw = 1;   // the window's width in 64-tile rows
  uint32_t spins = 0;
  for (;;) {
    uint32_t v[Q];
    uint32_t jn = NONE;
    bool failed = false;
    for (;;) {  // poll until every tile up to the nearest inclusive count is ready
      bool ready[Q], incl[Q];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        // rows outside the window count as ready, hold no count and are
        // never inclusive; before tile 0 lies an inclusive count of 0
        const int64_t idx = look - 1 - lane - 64 * q;
        const bool in = (uint32_t)q < qw;
        uint64_t st = 0;
        if (in && idx >= 0) st = __hip_atomic_load(&stat[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t fl = (uint32_t)(st >> 62);
        const bool rd = idx < 0 || (fl != 0 && (uint32_t)((st >> 32) & 0x3FFFFFFFu) == tag);
        ready[q] = !in || rd;
        incl[q] = in && (idx < 0 || (rd && fl == TILE_INCL));
        v[q] = (in && idx >= 0) ? (uint32_t)st : 0u;
      }
      jn = NONE;
#pragma unroll
      for (int q = Q - 1; q >= 0; --q)
        if (incl[q]) jn = (uint32_t)(lane + 64 * q);
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)jn, o, 64);
        jn = y < jn ? y : jn;
      }
      bool ok = true;
#pragma unroll
      for (int q = 0; q < Q; ++q)
        if ((uint32_t)(lane + 64 * q) <= jn && !ready[q]) ok = false;
      if (__all(ok)) break;
      if (++spins > LOOKBACK_SPIN_LIMIT) {
        if (lane == 0) atomicExch(&sc->lookbackFail, 1u);
        failed = true;
        break;
      }
      for (uint32_t z = 0; z < lb.sleep; ++z) __builtin_amdgcn_s_sleep(2);
    }
    if (failed) break;
    uint32_t x = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q)
      if ((uint32_t)(lane + 64 * q) <= jn) x += v[q];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
    excl += x;
    if (jn != NONE) break;  // reached an inclusive count
    look -= 64 * (int64_t)qw;
    qw = lb.q;
  }
  if (lane == 0)
    __hip_atomic_store(&stat[t], tileWord(TILE_INCL, tag, excl + agg), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

template <class At, int MODE, int D, bool STAR, int TW>
__global__ __launch_bounds__(256) void kBandOp(At at, const uint64_t *A, const uint64_t *B, Frame f, GridInfo g,
                                               uint32_t nw, BandBufs out, uint64_t *stat, DevScalars *sc,
                                               const uint32_t *inPre, uint32_t inBase, Lookback lb) {
  static_assert(TW == 32 || TW == 64, "wave 0 builds one word per lane");
  __shared__ uint64_t cw[TW], kw[TW];
  __shared__ uint32_t cpre[TW + 1], kpre[TW + 1];
  __shared__ double stage[TW * 64];
  __shared__ double wlast[TPB / 64];
  __shared__ uint32_t wcnt[TPB / 64];
  __shared__ uint32_t exclS;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint32_t tag = launchTag(sc->epoch);
  const bool skip = sc->overflow != 0;  // an earlier kernel of the step overflowed: no tile is claimed
  const uint32_t nTiles = (nw + TW - 1) / TW;
  const uint32_t base = (uint32_t)f.base;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t t = blockIdx.x; !skip && t < nTiles; t += nTiles) {  // one tile per workgroup
    const uint32_t w0 = t * TW;
    if (wid == 0) {
      const uint32_t w = w0 + (uint32_t)lane;
      uint64_t C = 0;
      if (lane < TW && w < nw) {
        if (MODE == CAND_STAR) C = candStarWord<D>(A, f, g, (int64_t)w);
        else if (MODE == CAND_UNION) C = A[w] | B[w];
        else C = A[w];
      }
      const uint32_t pc = (uint32_t)__builtin_popcountll(C);
      uint32_t x = pc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane < TW) {
        cw[lane] = C;
        cpre[lane] = x - pc;
        kw[lane] = 0;
      }
      if (lane == TW - 1) cpre[TW] = x;
    }
    __syncthreads();
    const uint32_t nc = cpre[TW];
    // slot 0: the grid point before the first candidate (tile 0: the
    // grid-minimum entry outside the frame, or nothing when the grid minimum
    // is the first candidate); slot s >= 1: candidate s - 1
    const bool hasPred = (t > 0) || base;
    // tile 0 always evaluates the grid-minimum entry when it lies outside
    // the frame (output entry 0), even without candidates of its own
    const uint32_t nslots = (nc > 0 || (t == 0 && base)) ? nc + 1 : 0;
    uint32_t kept = 0;
    double carry = 0.0;
    for (uint32_t s0 = 0; s0 < nslots; s0 += TPB) {
      const uint32_t s = s0 + threadIdx.x;
      uint32_t p;
      bool eval, cand;
      int j, bit;
      tileSlot<TW>(f, t, w0, cw, cpre, nc, s, p, eval, cand, j, bit);
      const double v = eval ? at(p) : 0.0;
      if (s == 0 && t == 0 && base) {  // output entry 0: the grid-minimum entry
        out.val[0] = v;
        out.pos[0] = POS_NONE;
      }
      const double up = __shfl_up(v, 1, 64);
      if (lane == 63) wlast[wid] = v;
      __syncthreads();
      const double prev = (lane > 0) ? up : (wid > 0 ? wlast[wid - 1] : carry);
      const double nextCarry = wlast[TPB / 64 - 1];
      bool keep = false;
      if (cand) keep = (s == 1 && !hasPred) || !isUndef(v) || v != prev;
      const uint64_t m = (uint64_t)__ballot(keep);
      if (lane == 0) wcnt[wid] = (uint32_t)__builtin_popcountll(m);
      __syncthreads();
      uint32_t woff = 0, tot = 0;
#pragma unroll
      for (int q = 0; q < TPB / 64; ++q) {
        const uint32_t cq = wcnt[q];
        if (q < wid) woff += cq;
        tot += cq;
      }
      if (keep) {
        stage[kept + woff + (uint32_t)__builtin_popcountll(m & below)] = v;
        atomicOr((unsigned long long *)&kw[j], 1ull << bit);
      }
      kept += tot;
      carry = nextCarry;
      __syncthreads();
    }
    if (wid == 0) {
      const uint64_t K = (lane < TW) ? kw[lane] : 0ull;
      const uint32_t pc = (uint32_t)__builtin_popcountll(K);
      uint32_t x = pc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane < TW) kpre[lane] = x - pc;
      if (lane == TW - 1) kpre[TW] = x;
      const uint32_t e = lookBack(stat, t, kept, tag, sc, lb);
      if (lane == 0) exclS = e;
    }
    __syncthreads();
    const uint32_t P = exclS;
    const uint64_t o0 = (uint64_t)base + P;
    if (o0 + kept <= out.cap) {
      for (uint32_t k = threadIdx.x; k < kept; k += TPB) {
        const int jj = wordOfItem<TW>(kpre, k);
        const int b = selectBit(kw[jj], (int)(k - kpre[jj]));
        out.val[o0 + k] = stage[k];
        out.pos[o0 + k] = ((w0 + (uint32_t)jj) << 6) | (uint32_t)b;
      }
    } else if (threadIdx.x == 0) {
      sc->overflow = 1;
    }
    if (threadIdx.x < TW) {
      const uint32_t w = w0 + threadIdx.x;
      if (w < nw) {
        out.bits[w] = kw[threadIdx.x];
        out.pre[w] = P + kpre[threadIdx.x];
      }
    }
    if (t == nTiles - 1 && threadIdx.x == 0) {
      out.bits[nw] = 0;
      out.pre[nw] = P + kept;
      uint64_t n = (uint64_t)base + P + kept;
      if (n > out.cap) {
        sc->overflow = 1;
        n = out.cap;
      }
      *out.n = (uint32_t)n;
      // algorithmic bytes of a star operation: the input band's values and
      // positions read once, the output band's written once, and both
      // bitmaps with their prefix counts
      if (STAR) sc->starBytes += 12.0 * ((double)inPre[nw] + inBase + (double)n) + 24.0 * (double)nw;
      // every tile has published its count and every workgroup has read the
      // launch tag: the next launch tags its status words anew
      __hip_atomic_store(&sc->epoch, sc->epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}

// Batches: the same single pass over the segmented bitmap of many level
// sets.  A tile's items are, per word, the grid-minimum entry of the level
// set that starts at the word (when its frame does not hold the grid
// minimum), then the word's candidates.  The first item of a level set is
// always kept; an in-frame entry of level set k goes to index
// baseIncl(k) + (in-frame entries before it), its grid-minimum entry just
// before its first in-frame entry.  The look-back counts in-frame entries.
template <int D> struct AtStarSeg {
  BandView all;
  const SegInfo *segs;
  OpParams op;
  __device__ double operator()(uint32_t p, uint32_t k) const { return AtStar<D>{segView(all, segs[k]), op}(p); }
};
template <int D> struct AtSelfSeg {
  BandView all;
  const SegInfo *segs;
  OpParams op;  // dt: one per level set
  __device__ double operator()(uint32_t p, uint32_t k) const {
    OpParams o = op;
    o.dt = op.dt + k;
    return AtSelf<D>{segView(all, segs[k]), o}(p);
  }
};
struct AtBoolSeg {
  BandView A, B;
  const SegInfo *segs;
  int kind;
  double wT, wS;
  __device__ double operator()(uint32_t p, uint32_t k) const {
    const SegInfo &S = segs[k];
    return AtBool{segView(A, S), segView(B, S), kind, wT, wS, nullptr}(p);
  }
};

static constexpr int TW_ITEMS = TWS * 64 + TWS;  // candidates and grid-minimum entries of a tile
static constexpr uint32_t NO_SEG = 0xFFFFFFFFu;

template <class At, int MODE, int D>
__global__ __launch_bounds__(256) void kBandOpSeg(At at, const uint64_t *A, const uint64_t *B, SegArgs sa,
                                                  uint32_t nw, BandBufs out, uint64_t *stat, DevScalars *sc,
                                                  Lookback lb) {
  constexpr int TW = TWS;
  __shared__ uint64_t cw[TW], kw[TW];
  __shared__ uint32_t ipre[TW + 1], kpre[TW + 1], wk[TW];
  __shared__ uint32_t extraMask;
  __shared__ double sval[TW_ITEMS];
  __shared__ uint32_t soff[TW_ITEMS], spos[TW_ITEMS];
  __shared__ uint16_t sseg[TW_ITEMS];
  __shared__ double wlast[TPB / 64];
  __shared__ uint32_t wlastSeg[TPB / 64];
  __shared__ uint32_t wcnt[TPB / 64], wcntIn[TPB / 64];
  __shared__ uint32_t exclS;
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint32_t tag = launchTag(sc->epoch);
  const bool skip = sc->overflow != 0;
  const uint32_t nTiles = (nw + TW - 1) / TW;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (uint32_t t = blockIdx.x; !skip && t < nTiles; t += nTiles) {  // one tile per workgroup
    const uint32_t w0 = t * TW;
    if (wid == 0) {
      const uint32_t w = w0 + (uint32_t)lane;
      uint64_t C = 0;
      uint32_t k = 0, extra = 0;
      if (lane < TW && w < nw) {
        k = sa.wseg[w];
        const SegInfo &S = sa.segs[k];
        if (MODE == CAND_STAR) C = candStarWord<D>(A + S.wordOff, S.f, S.g, (int64_t)(w - S.wordOff));
        else if (MODE == CAND_UNION) C = A[w] | B[w];
        else C = A[w];
        extra = (w == S.wordOff && S.f.base) ? 1u : 0u;
      }
      const uint32_t pc = (uint32_t)__builtin_popcountll(C) + extra;
      uint32_t x = pc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
      }
      const uint64_t em = (uint64_t)__ballot(extra != 0);
      if (lane < TW) {
        cw[lane] = C;
        ipre[lane] = x - pc;
        kw[lane] = 0;
        wk[lane] = k;
      }
      if (lane == TW - 1) ipre[TW] = x;
      if (lane == 0) extraMask = (uint32_t)em;
    }
    __syncthreads();
    const uint32_t ni = ipre[TW];
    uint32_t kept = 0, keptIn = 0;
    double carry = 0.0;
    uint32_t carrySeg = NO_SEG;
    for (uint32_t s0 = 0; ni > 0 && s0 < ni + 1; s0 += TPB) {
      const uint32_t s = s0 + threadIdx.x;
      uint32_t p = POS_NONE, k = NO_SEG;
      bool eval = false, item = false, isBase = false;
      int j = 0, bit = 0;
      if (s == 0) {  // the grid point before the first item, in its level set
        const int j0 = wordOfItem<TW>(ipre, 0);
        if (!((extraMask >> j0) & 1u)) {
          const uint32_t kk = wk[j0];
          const uint32_t lw = w0 + (uint32_t)j0 - sa.segs[kk].wordOff;
          const uint32_t p1 = (lw << 6) | (uint32_t)__builtin_ctzll(cw[j0]);
          if (p1 > 0) {
            p = realPred(sa.segs[kk].f, p1);
            k = kk;
            eval = true;
          }
        }
      } else if (s <= ni) {
        const uint32_t c = s - 1;
        j = wordOfItem<TW>(ipre, c);
        k = wk[j];
        const uint32_t ex = (extraMask >> j) & 1u;
        const uint32_t r = c - ipre[j];
        item = eval = true;
        if (ex && r == 0) {
          isBase = true;
        } else {
          bit = selectBit(cw[j], (int)(r - ex));
          p = ((w0 + (uint32_t)j - sa.segs[k].wordOff) << 6) | (uint32_t)bit;
        }
      }
      const double v = eval ? at(p, k) : 0.0;
      const double up = __shfl_up(v, 1, 64);
      const uint32_t upSeg = (uint32_t)__shfl_up((int)k, 1, 64);
      if (lane == 63) {
        wlast[wid] = v;
        wlastSeg[wid] = k;
      }
      __syncthreads();
      const double prev = (lane > 0) ? up : (wid > 0 ? wlast[wid - 1] : carry);
      const uint32_t prevSeg = (lane > 0) ? upSeg : (wid > 0 ? wlastSeg[wid - 1] : carrySeg);
      const double nextCarry = wlast[TPB / 64 - 1];
      const uint32_t nextCarrySeg = wlastSeg[TPB / 64 - 1];
      bool keep = false;
      if (item) keep = isBase || prevSeg != k || !isUndef(v) || v != prev;
      const bool keepIn = keep && !isBase;
      const uint64_t m = (uint64_t)__ballot(keep);
      const uint64_t mIn = (uint64_t)__ballot(keepIn);
      if (lane == 0) {
        wcnt[wid] = (uint32_t)__builtin_popcountll(m);
        wcntIn[wid] = (uint32_t)__builtin_popcountll(mIn);
      }
      __syncthreads();
      uint32_t woff = 0, tot = 0, woffIn = 0, totIn = 0;
#pragma unroll
      for (int q = 0; q < TPB / 64; ++q) {
        const uint32_t cq = wcnt[q], cqIn = wcntIn[q];
        if (q < wid) {
          woff += cq;
          woffIn += cqIn;
        }
        tot += cq;
        totIn += cqIn;
      }
      if (keep) {
        const uint32_t si = kept + woff + (uint32_t)__builtin_popcountll(m & below);
        const uint32_t inBefore = keptIn + woffIn + (uint32_t)__builtin_popcountll(mIn & below);
        sval[si] = v;
        soff[si] = sa.segs[k].baseIncl - (isBase ? 1u : 0u) + inBefore;
        spos[si] = p;
        sseg[si] = (uint16_t)k;
        if (!isBase) atomicOr((unsigned long long *)&kw[j], 1ull << bit);
      }
      kept += tot;
      keptIn += totIn;
      carry = nextCarry;
      carrySeg = nextCarrySeg;
      __syncthreads();
    }
    if (wid == 0) {
      const uint64_t K = (lane < TW) ? kw[lane] : 0ull;
      const uint32_t pc = (uint32_t)__builtin_popcountll(K);
      uint32_t x = pc;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane < TW) kpre[lane] = x - pc;
      if (lane == TW - 1) kpre[TW] = x;
      const uint32_t e = lookBack(stat, t, keptIn, tag, sc, lb);
      if (lane == 0) exclS = e;
    }
    __syncthreads();
    const uint32_t P = exclS;
    bool over = false;
    for (uint32_t q = threadIdx.x; q < kept; q += TPB) {
      const uint64_t o = (uint64_t)P + soff[q];
      if (o < out.cap) {
        out.val[o] = sval[q];
        out.pos[o] = spos[q];
        out.seg[o] = sseg[q];
      } else {
        over = true;
      }
    }
    if (over) sc->overflow = 1;
    if (threadIdx.x < TW) {
      const uint32_t w = w0 + threadIdx.x;
      if (w < nw) {
        out.bits[w] = kw[threadIdx.x];
        out.pre[w] = P + kpre[threadIdx.x];
      }
    }
    if (t == nTiles - 1 && threadIdx.x == 0) {
      out.bits[nw] = 0;
      out.pre[nw] = P + keptIn;
      uint64_t n = (uint64_t)sa.totalBase + P + keptIn;
      if (n > out.cap) {
        sc->overflow = 1;
        n = out.cap;
      }
      *out.n = (uint32_t)n;
      __hip_atomic_store(&sc->epoch, sc->epoch + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
  }
}


// Single-pass band operations: off by default (the multi-kernel pipeline
// below is faster at the measured sizes); LSG_FUSED_BANDOPS=1 or
// setSinglePassBandOps turns them on; a look-back wait that runs out turns
// them off for the rest of the process.  LSG_LB_WINDOW (1..8 rows of 64
// tiles), LSG_LB_SLEEP (polls' backoff) and LSG_BAND_TILE_WORDS (32 or 64)
// tune them.
static uint32_t envU32(const char *name, uint32_t dflt, uint32_t lo, uint32_t hi) {
  const char *e = getenv(name);
  if (!e || !e[0]) return dflt;
  const unsigned long v = strtoul(e, nullptr, 10);
  return (uint32_t)std::min<unsigned long>(std::max<unsigned long>(v, lo), hi);
}
static bool g_fused = envU32("LSG_FUSED_BANDOPS", 0, 0, 1) != 0;
static const Lookback g_lb{envU32("LSG_LB_WINDOW", LOOKBACK_PER_LANE, 1, LOOKBACK_PER_LANE),
                           envU32("LSG_LB_SLEEP", 1, 1, 64)};
static const int g_bandTW = envU32("LSG_BAND_TILE_WORDS", 32, 32, 64) >= 64 ? 64 : 32;
bool singlePassBandOps() { return g_fused; }
void disableSinglePassBandOps() { g_fused = false; }
void setSinglePassBandOps(bool on) { g_fused = on; }

template <class At, int MODE, int D, bool STAR>
static void bandOp(hipStream_t s, const At &at, const uint64_t *A, const uint64_t *B, const Frame &f,
                   const GridInfo &g, BandBufs &out, uint64_t *stat, DevScalars *sc, const BandBufs &in) {
  const uint32_t nw = (uint32_t)f.nwords;
  if (g_bandTW == 64)
    kBandOp<At, MODE, D, STAR, 64><<<(nw + 63) / 64, TPB, 0, s>>>(at, A, B, f, g, nw, out, stat, sc, in.pre,
                                                                  (uint32_t)f.base, g_lb);
  else
    kBandOp<At, MODE, D, STAR, 32><<<(nw + 31) / 32, TPB, 0, s>>>(at, A, B, f, g, nw, out, stat, sc, in.pre,
                                                                  (uint32_t)f.base, g_lb);
}

void launchStarOpScratch(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f,
                         const GridInfo &g, const OpParams &op, uint64_t *stat, DevScalars *sc) {
  const BandView v = viewOf(in, f, g);
  if (D == 2) bandOp<AtStar<2>, CAND_STAR, 2, true>(s, AtStar<2>{v, op}, in.bits, nullptr, f, g, out, stat, sc, in);
  else bandOp<AtStar<3>, CAND_STAR, 3, true>(s, AtStar<3>{v, op}, in.bits, nullptr, f, g, out, stat, sc, in);
}

__global__ void kClearFlags(DevScalars *sc) {
  sc->overflow = 0;
  sc->lookbackFail = 0;
  sc->reframe = 0;
}
void launchClearFlags(hipStream_t s, DevScalars *sc) { kClearFlags<<<1, 1, 0, s>>>(sc); }

// ------------------------------------------------------- host wrappers ----

static SegArgs noSeg() { return SegArgs{nullptr, 0, 0, nullptr, nullptr, nullptr}; }
static SegArgs segArgs(const SegCtx &c, const uint16_t *iseg, uint16_t *oseg) {
  return SegArgs{c.segs, c.nseg, c.totalBase, c.wseg, iseg, oseg};
}

static void recordEvent(hipStream_t s, hipEvent_t e, bool external) {
  if (!e) return;
  if (external) (void)hipEventRecordWithFlags(e, s, hipEventRecordExternal);
  else (void)hipEventRecord(e, s);
}

template <class Ev, bool SEG>
static void compactInto(hipStream_t s, const Ev &ev, const Workspace &ws, const uint32_t *ncandPtr,
                        const uint32_t *cpos, const uint32_t *cpre, const uint64_t *cwords, uint32_t nw, int base,
                        BandBufs &out, DevScalars *sc, const SegArgs &sa, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr, bool externalEvents = false) {
  recordEvent(s, e0, externalEvents);
  kEvalKeep<Ev, SEG><<<g_scanBlocks, TPB, 0, s>>>(ev, ncandPtr, ws.out, ws.partials, sc, sa.iseg);
  recordEvent(s, e1, externalEvents);
  kKeepScatter<SEG><<<g_scanBlocks, TPB, 0, s>>>(ws.out, cpos, ncandPtr, ws.partials, ws.ks, ws.keepBits, out.val,
                                                 out.pos, out.n, out.cap, sc, sa.iseg, out.seg);
  kBuildIndex<SEG><<<gridFor((size_t)nw + 1), TPB, 0, s>>>(cwords, cpre, ws.ks, ws.keepBits, nw, base, ncandPtr,
                                                           out.bits, out.pre, sc, sa);
}

template <int D, int MODE, bool SEG>
static void candidates(hipStream_t s, const uint64_t *A, const uint64_t *B, const Frame &f, const GridInfo &g,
                       uint32_t nw, Workspace &ws, DevScalars *sc, const SegArgs &sa) {
  kCandCount<D, MODE, SEG><<<g_scanBlocks, TPB, 0, s>>>(A, B, f, g, nw, ws.cbits, ws.partials, sc, sa);
  kCandList<SEG><<<g_scanBlocks, TPB, 0, s>>>(ws.cbits, nw, f.base, ws.partials, ws.cpre, ws.cpos, ws.ccap, ws.ncand,
                                              sc, sa, ws.cseg);
}

template <int D, bool SEG>
static void starOp(hipStream_t s, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g, uint32_t nw,
                   Workspace &ws, const OpParams &op, DevScalars *sc, const SegArgs &sa0, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr, bool externalEvents = false) {
  candidates<D, 0, SEG>(s, in.bits, nullptr, f, g, nw, ws, sc, sa0);
  SegArgs sa = sa0;
  sa.iseg = ws.cseg;
  sa.oseg = out.seg;
  EvalStar<D, SEG> ev{viewOf(in, f, g), sa, ws.cpos, op};
  compactInto<EvalStar<D, SEG>, SEG>(s, ev, ws, ws.ncand, ws.cpos, ws.cpre, ws.cbits, nw, f.base, out, sc, sa, e0, e1,
                                     externalEvents);
}

// Two Expand cycles in one band operation: the candidates are the star
// dilation of the star dilation of the input's run starts (a superset of
// the second cycle's visits: the first cycle's run starts are among its
// candidates), each evaluated with expandTwice.
template <int D>
static void starOp2(hipStream_t s, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g, uint32_t nw,
                    Workspace &ws, const OpParams &op, DevScalars *sc, hipEvent_t e0, hipEvent_t e1,
                    bool externalEvents) {
  const SegArgs sa0 = noSeg();
  kCandCount<D, 0, false><<<g_scanBlocks, TPB, 0, s>>>(in.bits, nullptr, f, g, nw, ws.cbits2, ws.partials, sc, sa0);
  candidates<D, 0, false>(s, ws.cbits2, nullptr, f, g, nw, ws, sc, sa0);
  SegArgs sa = sa0;
  sa.iseg = ws.cseg;
  sa.oseg = out.seg;
  EvalStar2<D> ev{viewOf(in, f, g), ws.cpos, op};
  compactInto<EvalStar2<D>, false>(s, ev, ws, ws.ncand, ws.cpos, ws.cpre, ws.cbits, nw, f.base, out, sc, sa, e0, e1,
                                   externalEvents);
}

template <int D>
static void updRebuildOp(hipStream_t s, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                         uint32_t nw, Workspace &ws, const OpParams &op, DevScalars *sc) {
  const SegArgs sa0 = noSeg();
  candidates<D, 0, false>(s, in.bits, nullptr, f, g, nw, ws, sc, sa0);
  SegArgs sa = sa0;
  sa.iseg = ws.cseg;
  sa.oseg = out.seg;
  EvalUpdRebuild<D> ev{viewOf(in, f, g), ws.cpos, op};
  compactInto<EvalUpdRebuild<D>, false>(s, ev, ws, ws.ncand, ws.cpos, ws.cpre, ws.cbits, nw, f.base, out, sc, sa);
}

template <int D, bool SEG>
static void selfOp(hipStream_t s, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g, uint32_t nw,
                   Workspace &ws, const OpParams &op, DevScalars *sc, const SegArgs &sa0) {
  SegArgs sa = sa0;
  sa.iseg = in.seg;
  sa.oseg = out.seg;
  EvalSelf<D, SEG> ev{viewOf(in, f, g), sa, op};
  compactInto<EvalSelf<D, SEG>, SEG>(s, ev, ws, in.n, in.pos, in.pre, in.bits, nw, f.base, out, sc, sa);
}

template <int D, bool SEG>
static void boolOp(hipStream_t s, const BandBufs &A, const BandBufs &B, BandBufs &out, const Frame &f,
                   const GridInfo &g, uint32_t nw, Workspace &ws, int kind, double wT, double wS, DevScalars *sc,
                   const SegArgs &sa0, uint32_t *movedDown = nullptr) {
  candidates<D, 1, SEG>(s, A.bits, B.bits, f, g, nw, ws, sc, sa0);
  SegArgs sa = sa0;
  sa.iseg = ws.cseg;
  sa.oseg = out.seg;
  EvalBool<SEG> ev{viewOf(A, f, g), viewOf(B, f, g), sa, ws.cpos, kind, wT, wS, movedDown};
  compactInto<EvalBool<SEG>, SEG>(s, ev, ws, ws.ncand, ws.cpos, ws.cpre, ws.cbits, nw, f.base, out, sc, sa);
}

template <class At, int MODE, int D>
static void bandOpSeg(hipStream_t s, const At &at, const uint64_t *A, const uint64_t *B, const SegCtx &c,
                      BandBufs &out, uint64_t *stat, DevScalars *sc) {
  const uint32_t nTiles = (c.nw + TWS - 1) / TWS;
  kBandOpSeg<At, MODE, D><<<nTiles, TPB, 0, s>>>(at, A, B, segArgs(c, nullptr, nullptr), c.nw, out, stat, sc, g_lb);
}

void launchStarOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc, hipEvent_t e0, hipEvent_t e1,
                  bool externalEvents) {
  if (g_fused) {
    const BandView v = viewOf(in, f, g);
    recordEvent(s, e0, externalEvents);
    if (D == 2) bandOp<AtStar<2>, CAND_STAR, 2, true>(s, AtStar<2>{v, op}, in.bits, nullptr, f, g, out, ws.tileStat, sc, in);
    else bandOp<AtStar<3>, CAND_STAR, 3, true>(s, AtStar<3>{v, op}, in.bits, nullptr, f, g, out, ws.tileStat, sc, in);
    recordEvent(s, e1, externalEvents);
    return;
  }
  const uint32_t nw = (uint32_t)f.nwords;
  if (D == 2) starOp<2, false>(s, in, out, f, g, nw, ws, op, sc, noSeg(), e0, e1, externalEvents);
  else starOp<3, false>(s, in, out, f, g, nw, ws, op, sc, noSeg(), e0, e1, externalEvents);
}

void launchStarOp2(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                   Workspace &ws, const OpParams &op, DevScalars *sc, hipEvent_t e0, hipEvent_t e1,
                   bool externalEvents) {
  const uint32_t nw = (uint32_t)f.nwords;
  if (D == 2) starOp2<2>(s, in, out, f, g, nw, ws, op, sc, e0, e1, externalEvents);
  else starOp2<3>(s, in, out, f, g, nw, ws, op, sc, e0, e1, externalEvents);
}

void launchUpdRebuildOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                        Workspace &ws, const OpParams &op, DevScalars *sc) {
  const uint32_t nw = (uint32_t)f.nwords;
  if (D == 2) updRebuildOp<2>(s, in, out, f, g, nw, ws, op, sc);
  else updRebuildOp<3>(s, in, out, f, g, nw, ws, op, sc);
}

void launchSelfOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc) {
  if (g_fused) {
    const BandView v = viewOf(in, f, g);
    if (D == 2) bandOp<AtSelf<2>, CAND_SELF, 2, false>(s, AtSelf<2>{v, op}, in.bits, nullptr, f, g, out, ws.tileStat, sc, in);
    else bandOp<AtSelf<3>, CAND_SELF, 3, false>(s, AtSelf<3>{v, op}, in.bits, nullptr, f, g, out, ws.tileStat, sc, in);
    return;
  }
  const uint32_t nw = (uint32_t)f.nwords;
  if (D == 2) selfOp<2, false>(s, in, out, f, g, nw, ws, op, sc, noSeg());
  else selfOp<3, false>(s, in, out, f, g, nw, ws, op, sc, noSeg());
}

void launchBoolean(hipStream_t s, int D, const BandBufs &A, const BandBufs &B, BandBufs &out, const Frame &f,
                   const GridInfo &g, Workspace &ws, int kind, double wTarget, double wSource, DevScalars *sc,
                   uint32_t *movedDown) {
  if (g_fused) {
    const AtBool at{viewOf(A, f, g), viewOf(B, f, g), kind, wTarget, wSource, movedDown};
    if (D == 2) bandOp<AtBool, CAND_UNION, 2, false>(s, at, A.bits, B.bits, f, g, out, ws.tileStat, sc, A);
    else bandOp<AtBool, CAND_UNION, 3, false>(s, at, A.bits, B.bits, f, g, out, ws.tileStat, sc, A);
    return;
  }
  const uint32_t nw = (uint32_t)f.nwords;
  if (D == 2) boolOp<2, false>(s, A, B, out, f, g, nw, ws, kind, wTarget, wSource, sc, noSeg(), movedDown);
  else boolOp<3, false>(s, A, B, out, f, g, nw, ws, kind, wTarget, wSource, sc, noSeg(), movedDown);
}

// Batch variants: one launch per kernel for every level set of the batch.
// The frame and grid arguments of the kernels are unused (each item takes
// its level set's from the table).
void launchStarOpSeg(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const SegCtx &c, Workspace &ws,
                     const OpParams &op, DevScalars *sc) {
  if (g_fused) {
    const BandView all = viewOf(in, Frame{}, GridInfo{});
    if (D == 2) bandOpSeg<AtStarSeg<2>, CAND_STAR, 2>(s, AtStarSeg<2>{all, c.segs, op}, in.bits, nullptr, c, out,
                                                      ws.tileStat, sc);
    else bandOpSeg<AtStarSeg<3>, CAND_STAR, 3>(s, AtStarSeg<3>{all, c.segs, op}, in.bits, nullptr, c, out,
                                               ws.tileStat, sc);
    return;
  }
  Frame f{};
  GridInfo g{};
  if (D == 2) starOp<2, true>(s, in, out, f, g, c.nw, ws, op, sc, segArgs(c, nullptr, nullptr));
  else starOp<3, true>(s, in, out, f, g, c.nw, ws, op, sc, segArgs(c, nullptr, nullptr));
}

void launchSelfOpSeg(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const SegCtx &c, Workspace &ws,
                     const OpParams &op, DevScalars *sc) {
  if (g_fused) {
    const BandView all = viewOf(in, Frame{}, GridInfo{});
    if (D == 2) bandOpSeg<AtSelfSeg<2>, CAND_SELF, 2>(s, AtSelfSeg<2>{all, c.segs, op}, in.bits, nullptr, c, out,
                                                      ws.tileStat, sc);
    else bandOpSeg<AtSelfSeg<3>, CAND_SELF, 3>(s, AtSelfSeg<3>{all, c.segs, op}, in.bits, nullptr, c, out,
                                               ws.tileStat, sc);
    return;
  }
  Frame f{};
  GridInfo g{};
  if (D == 2) selfOp<2, true>(s, in, out, f, g, c.nw, ws, op, sc, segArgs(c, nullptr, nullptr));
  else selfOp<3, true>(s, in, out, f, g, c.nw, ws, op, sc, segArgs(c, nullptr, nullptr));
}

void launchBooleanSeg(hipStream_t s, int D, const BandBufs &A, const BandBufs &B, BandBufs &out, const SegCtx &c,
                      Workspace &ws, int kind, double wTarget, double wSource, DevScalars *sc) {
  if (g_fused) {
    const AtBoolSeg at{viewOf(A, Frame{}, GridInfo{}), viewOf(B, Frame{}, GridInfo{}), c.segs, kind, wTarget, wSource};
    if (D == 2) bandOpSeg<AtBoolSeg, CAND_UNION, 2>(s, at, A.bits, B.bits, c, out, ws.tileStat, sc);
    else bandOpSeg<AtBoolSeg, CAND_UNION, 3>(s, at, A.bits, B.bits, c, out, ws.tileStat, sc);
    return;
  }
  Frame f{};
  GridInfo g{};
  if (D == 2) boolOp<2, true>(s, A, B, out, f, g, c.nw, ws, kind, wTarget, wSource, sc, segArgs(c, nullptr, nullptr));
  else boolOp<3, true>(s, A, B, out, f, g, c.nw, ws, kind, wTarget, wSource, sc, segArgs(c, nullptr, nullptr));
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

// Wave-level compaction of the active entries (|phi| <= 0.5, defined) of a
// band: each wave scans 64 entries at a time, queues the active ones in LDS
// and runs `fn(i)` on full waves, so the expensive per-point work never runs
// with idle lanes.  The loop trip count is wave-uniform.
template <class Fn> __device__ inline void forEachActive(const double *val, uint32_t n, Fn &fn) {
  __shared__ uint32_t queue[TPB / 64][128];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  uint32_t *q = queue[wid];
  uint32_t cnt = 0;
  // the block's contiguous chunk of entries, 64-entry slices dealt to its waves
  uint32_t lo, hi;
  blockChunk(n, lo, hi);
  const uint32_t stride = (TPB / 64) * 64u;
  uint32_t base = lo + (uint32_t)wid * 64u;
  // One call site of fn, so that it is inlined once: the loop alternates
  // between queueing one 64-entry slice and draining a full wave of queued
  // entries; after the last slice the remainder (< 64) is drained.
  for (;;) {
    if (base < hi) {
      const uint32_t i = base + lane;
      bool act = false;
      if (i < hi) {
        const double v = val[i];
        act = !isUndef(v) && fabs(v) <= 0.5;
      }
      const uint64_t m = (uint64_t)__ballot(act ? 1 : 0);
      if (act) q[cnt + (uint32_t)__builtin_popcountll(m & ((1ull << lane) - 1ull))] = i;
      cnt += (uint32_t)__builtin_popcountll(m);
      base += stride;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    const bool last = base >= hi;
    if (cnt >= 64 || (last && cnt > 0)) {
      const uint32_t take = cnt < 64 ? cnt : 64u;
      if ((uint32_t)lane < take) fn(q[lane]);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const uint32_t rem = cnt - take;
      const uint32_t moved = ((uint32_t)lane < rem) ? q[64 + lane] : 0u;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      if ((uint32_t)lane < rem) q[lane] = moved;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      cnt = rem;
    }
    if (last && cnt == 0) break;
  }
}

// Active entries (|phi| <= 0.5, defined; lsAdvect.hpp:498-499) in entry
// order: per-block counts, then the list (two-kernel scan).  The rate kernel
// then gives every lane one active point.
__device__ inline bool isActive(double v) { return !isUndef(v) && fabs(v) <= 0.5; }

__global__ __launch_bounds__(256) void kActiveCount(const double *val, const uint32_t *nptr, uint32_t *partials,
                                                    DevScalars *sc, int resetRates, int keepAlphas) {
  LSG_ABORT_IF_OVERFLOW(sc);
  if (resetRates && logicalBlock() == 0 && threadIdx.x == 0) {  // the rate pass starts from zero maxima
    sc->maxAbsBits = 0;
    sc->sllfKey = 0;
    if (!keepAlphas) sc->alphaBits[0] = sc->alphaBits[1] = sc->alphaBits[2] = 0;
    sc->activeCount = 0;
  }
  const uint32_t n = *nptr;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t acc = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += TPB) acc += isActive(val[i]) ? 1u : 0u;
  const uint32_t t = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = t;
}

__global__ __launch_bounds__(256) void kActiveList(const double *val, const uint32_t *nptr, const uint32_t *partials,
                                                   uint32_t *list, uint32_t *nact, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t off = blockOffset(partials);
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t i = t + threadIdx.x;
    const bool a = (i < hi) && isActive(val[i]);
    uint32_t tot;
    uint64_t wm;
    const uint32_t p = off + blockBallotScan(a, tot, wm);
    if (a) list[p] = i;
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) *nact = off;
}

// Defined entries in entry order (point ids of the host-velocity path).
__global__ __launch_bounds__(256) void kDefinedCount(const double *val, const uint32_t *nptr, uint32_t *partials) {
  const uint32_t n = *nptr;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t acc = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += TPB) acc += isUndef(val[i]) ? 0u : 1u;
  const uint32_t t = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = t;
}
__global__ __launch_bounds__(256) void kDefinedList(const double *val, const uint32_t *nptr, const uint32_t *partials,
                                                    uint32_t *list, uint32_t *ndef) {
  const uint32_t n = *nptr;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t off = blockOffset(partials);
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t i = t + threadIdx.x;
    const bool a = (i < hi) && !isUndef(val[i]);
    uint32_t tot;
    uint64_t wm;
    const uint32_t p = off + blockBallotScan(a, tot, wm);
    if (a) list[p] = i;
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) *ndef = off;
}

// Values of the lower level sets and the top at grid point c (the lower
// ones share the top's frame); lv[lw.n] = phiTop.
__device__ inline void levelValuesAt(const BandView &b, const LowerViews &lw, const int32_t c[3], double phiTop,
                                     double *lv) {
  const BandView::Loc l = b.locate(c[0], c[1], c[2]);
  for (int m = 0; m < lw.n; ++m) {
    BandView v = b;
    v.bits = lw.bits[m];
    v.pre = lw.pre[m];
    v.val = lw.val[m];
    lv[m] = v.val[v.entryOf(l, v.bits[l.w], v.pre[l.w])];
  }
  lv[lw.n] = phiTop;
}
// Material of findGlobalAlphas (lsAdvect.hpp:133-141): the first level set,
// bottom up, whose value is <= value + wrappingLayerEpsilon.
__device__ inline int alphaMaterial(const double *lv, int L, double value) {
  for (int m = 0; m < L; ++m)
    if (lv[m] <= value + 1e-4) return m;
  return L - 1;
}

LSG_HD inline int hostRecordsOf(int scheme, int D) {
  if (scheme == S_LF) return 2;
  if (scheme == S_LLF || scheme == S_LLFA || scheme == S_SLLF) return 1 + ((D == 3) ? 27 : 9);
  return 1;
}
int hostVelocityRecords(int spatial, int D) { return hostRecordsOf(schemeIdOf(spatial), D); }

// Inputs of VelocityField::getScalarVelocity / getVectorVelocity for every
// active point: coordinate index * gridDelta, the scheme's normal, the point
// id (rank among the defined points of the expanded band) and -1 (called
// for every material) -- the scheme's call (lsEngquistOsher.hpp:140-145,
// lsWENO.hpp:132-137, lsLaxFriedrichs.hpp:136-141,
// lsLocalLaxFriedrichs.hpp:157-162); for LF then findGlobalAlphas' call
// with its own normal and the point's material (lsAdvect.hpp:145-171); for
// LLF then the 3^D alpha box, coordinate + offset * gridDelta with the box
// point's normal (lsLocalLaxFriedrichs.hpp:193-222); for the analytical LLF
// the 3^D box with the central differences getDissipationAlpha receives
// (lsLocalLaxFriedrichsAnalytical.hpp:191-214); for the stencil local
// Lax-Friedrichs scheme the 3^D box with the box points' first-order normals
// (lsStencilLocalLaxFriedrichsScalar.hpp:276-298).  8 doubles per record.
template <int D>
__global__ void kVelInputs(BandView b, LowerViews lw, const uint32_t *alist, const uint32_t *nact,
                           const uint32_t *dlist, const uint32_t *ndef, RateParams rp, double *out) {
  const uint32_t n = *nact, nd = *ndef;
  const int scheme = schemeIdOf(rp.spatial);
  const int R = hostRecordsOf(scheme, D);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t i = alist[j];
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    const double phi0 = b.val[i];
    double nrm[3];
    schemeNormal<D>(b, c, phi0, scheme, rp.order, rp.calcNormals != 0, nrm);
    uint32_t lo = 0, hi = nd;  // first defined entry >= i
    while (lo < hi) {
      const uint32_t m = (lo + hi) >> 1;
      if (dlist[m] < i) lo = m + 1;
      else hi = m;
    }
    double coord[3];
    for (int d = 0; d < 3; ++d) coord[d] = (d < D) ? (double)c[d] * b.g.dx : 0.0;
    double *o = out + 8 * (size_t)j * R;
    for (int d = 0; d < 3; ++d) o[d] = coord[d];
    for (int d = 0; d < 3; ++d) o[3 + d] = nrm[d];
    o[6] = (double)lo;
    o[7] = -1.0;
    if (scheme == S_LF) {
      double gn[3];
      globalAlphaNormal<D>(b, c, gn);
      double lv[MAX_LEVEL_SETS];
      levelValuesAt(b, lw, c, phi0, lv);
      double *r = o + 8;
      for (int d = 0; d < 3; ++d) r[d] = coord[d];
      for (int d = 0; d < 3; ++d) r[3 + d] = gn[d];
      r[6] = (double)lo;
      r[7] = (double)alphaMaterial(lv, lw.n + 1, phi0);
    } else if (scheme == S_LLF || scheme == S_LLFA || scheme == S_SLLF) {
      double box[(D == 3) ? 27 : 9];
      if (scheme == S_SLLF) gatherBox<D>(b, c, box);
      for (int q = 0; q < R - 1; ++q) {
        const int ni[3] = {q % 3 - 1, (q / 3) % 3 - 1, q / 9 - 1};
        double bn[3];
        if (scheme == S_LLF) llfBoxNormal<D>(b, c, q, bn);
        else if (scheme == S_LLFA) llfaBoxDiffs<D>(b, c, q, bn);
        else sllfNormal<D>(box, b.g.dx, ni[0], ni[1], D == 3 ? ni[2] : 0, bn);
        double *r = o + 8 * (1 + q);
        for (int d = 0; d < 3; ++d) r[d] = (d < D) ? coord[d] + (double)ni[d] * b.g.dx : 0.0;
        for (int d = 0; d < 3; ++d) r[3 + d] = bn[d];
        r[6] = (double)lo;
        r[7] = -1.0;
      }
    }
  }
}

// Second round of the stencil local Lax-Friedrichs host calls: for active
// point j, material m and box point q, the scalar velocity at the box normal
// moved by -DN and +DN along each axis k (lsStencilLocalLaxFriedrichsScalar.
// hpp:318-338), DN from the box point's velocity of the first round (the
// simple adaptive epsilon, :62-74), the axis moves replayed in the
// reference's order (each undone before the next axis).  Record
// ((j L + m) 3^D + q) 2 D + 2 k + s, 8 doubles: coordinate, normal, point id,
// material m -- or -2 where the reference makes no call: the point's scalar
// velocity is 0 (:264-266) or an earlier box point's normal was corrupted
// (:283-288).  `in` and `vel` are the first round's records and results.
template <int D>
__global__ void kSllfInputs(const uint32_t *nact, int L, const double *in, const double *vel, double baseEps,
                            double *out) {
  constexpr int NB = (D == 3) ? 27 : 9, R = 1 + NB;
  const uint32_t n = *nact;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n * (uint32_t)L; t += gridDim.x * blockDim.x) {
    const uint32_t j = t / (uint32_t)L;
    const int m = (int)(t % (uint32_t)L);
    const double *c0 = in + 8 * (size_t)j * R;
    const double *h0 = vel + ((size_t)j * R * L + m) * 4;
    double sv = h0[0];
    for (int i = 0; i < D; ++i) sv += h0[1 + i] * c0[3 + i];
    bool alive = sv != 0.;
    double *o = out + 8 * (size_t)t * NB * 2 * D;  // 8 doubles per record
    for (int q = 0; q < NB; ++q) {
      const double *r = c0 + 8 * (1 + q);
      const double *h = vel + (((size_t)j * R + 1 + q) * L + m) * 4;
      const double ln[3] = {r[3], r[4], r[5]};
      alive = alive && !sllfCorrupt(ln);
      double lsv = h[0];
      for (int d = 0; d < D; ++d) lsv += h[1 + d] * ln[d];
      const double DN = sllfEpsilon(baseEps, lsv);
      double np[3] = {ln[0], ln[1], ln[2]}, nn[3] = {ln[0], ln[1], ln[2]};
      for (int k = 0; k < D; ++k) {
        np[k] -= DN;
        nn[k] += DN;
        for (int sgn = 0; sgn < 2; ++sgn) {
          double *w = o + 8 * ((q * D + k) * 2 + sgn);
          const double *nv = sgn ? nn : np;
          for (int d = 0; d < 3; ++d) w[d] = r[d];
          for (int d = 0; d < 3; ++d) w[3 + d] = nv[d];
          w[6] = r[6];
          w[7] = alive ? (double)m : -2.0;
        }
        np[k] += DN;
        nn[k] -= DN;
      }
    }
  }
}

void launchSllfInputs(hipStream_t s, int D, const uint32_t *nact, int L, const double *in, const double *vel,
                      double baseEps, double *out) {
  if (D == 2) kSllfInputs<2><<<rateBlocks(), TPB, 0, s>>>(nact, L, in, vel, baseEps, out);
  else kSllfInputs<3><<<rateBlocks(), TPB, 0, s>>>(nact, L, in, vel, baseEps, out);
}

// Hamiltonian + dissipation per active entry of the expanded band
// (Advect::integrateTime, lsAdvect.hpp:495-593, one level set): writes the
// stage velocity and reduces max |H - D| (and the LLF/LLLF alphas).
template <int D, int SCHEME, bool SEG> struct RatePointFn {
  BandView b;
  RateParams rp;
  Vel vel;
  double *rates;
  const float *mlp;  // learned weno5 parameters staged in LDS
  float *io;         // learned weno5: this lane's network inputs/outputs (LDS)
  double ga[3];
  double la[3];
  double lmax;
  unsigned long long act;
  // batch: level set of each entry, per-level-set maxima; the running
  // maximum of a lane belongs to curSeg and is flushed when that changes
  SegArgs sa;
  unsigned long long *segMax;
  uint32_t curSeg;
  unsigned long long sllfKey;  // SLLF: key of this lane's last active point that set its alphas
  int learnedPhase;            // learned weno5: 2 when the wave has run the network (waveMlp)
  double *stc;                 // learned weno5: this lane's axis values (LDS, stride 64)
  // learned weno5, first phase: this lane's network inputs for entry i, and
  // its axis values kept for the second phase
  __device__ __attribute__((always_inline)) void prepareLearned(uint32_t i) {
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    double sts[D][7];
    gatherAxes<D, 3>(b, c, sts);
    const double phi0 = b.val[i];
    LSG_UNROLL
    for (int d = 0; d < D; ++d) sts[d][3] = phi0;
    LSG_UNROLL
    for (int d = 0; d < D; ++d)
      LSG_UNROLL
      for (int k = 0; k < 7; ++k) stc[(d * 7 + k) * 64] = sts[d][k];
    double gm[D], gp[D];
    weno5LearnedAxes<D>(sts, b.g.dx, mlp, io, gm, gp, 1);
  }
  __device__ __attribute__((always_inline)) void operator()(uint32_t i, uint32_t j) {
    ++act;
    BandView v = b;
    Vel ve = vel;
    if (!SEG && rp.hostVel) {  // velocities evaluated on the host for active point j
      const double *h = rp.hostVel + 4 * (size_t)j * rp.hostRec;
      ve.kind = 0;
      ve.b = h[0];
      ve.vec[0] = h[1];
      ve.vec[1] = h[2];
      ve.vec[2] = h[3];
      if (SCHEME == S_LLF || SCHEME == S_LLFA || SCHEME == S_SLLF) {  // the alpha box follows the scheme's call
        ve.box = h + 4;
        ve.boxStride = 4;
      }
      if (SCHEME == S_SLLF) ve.sllf = rp.hostSllf + (size_t)j * ((D == 3) ? 27 : 9) * 2 * D;
    }
    if (SEG) {
      const uint32_t k = sa.iseg[i];
      const SegInfo &S = sa.segs[k];
      v = segView(b, S);
      ve = S.vel;
      if (k != curSeg) {
        if (lmax > 0.) atomicMax(&segMax[curSeg], (unsigned long long)__double_as_longlong(lmax));
        lmax = 0.;
        if (SCHEME == S_LLLF || SCHEME == S_LLF) {  // the level set's local alpha maxima
          for (int d = 0; d < D; ++d) {
            if (la[d] > 0.) atomicMax(&rp.segAlpha[3 * curSeg + d], (unsigned long long)__double_as_longlong(la[d]));
            la[d] = 0.;
          }
        }
        if (SCHEME == S_LF)  // the level set's global alphas
          for (int d = 0; d < 3; ++d) ga[d] = __longlong_as_double((long long)rp.segGA[3 * k + d]);
        curSeg = k;
      }
    }
    if (!SEG && rp.voidMarks && rp.voidMarks[rp.voidRank[i]] != 0.) {  // void point: velocity 0
      rates[i] = 0.;
      if (rp.dissOut) rp.dissOut[i] = 0.;
      return;
    }
    int32_t c[3];
    candCoords(v, i, v.pos[i], c);
    double velocity;
    bool used = false;
    double dn = 0.;
    rates[i] = ratePoint<D, SCHEME>(v, c, v.val[i], rp, ve, ga, la, velocity, mlp, io, &used, &dn, learnedPhase, stc,
                                    64);
    if (!SEG && rp.dissOut) rp.dissOut[i] = dn;
    if (SCHEME == S_SLLF && used) sllfKey = (unsigned long long)(j + 1) << 4;  // lanes visit j in increasing order
    if (velocity != 0.) {
      const double a = fabs(velocity);
      if (lmax < a) lmax = a;
    }
  }
};

// ------------------------------------------- learned weno5 on the matrix cores
//
// The network evaluations of a wave (64 lanes x 2D, one row of 5 inputs each
// in LDS, zeros where a lane does not use the network) in blocks of 32
// rows, every layer a v_mfma_f32_32x32x16_bf16 on the transposed
// activations: D[m][n] = sum_k A[m][k] B[k][n] with m = unit, n = sample.
// Accumulator element r of lane l holds unit U(h, r) = 8(r/4) + 4h + r%4
// (h = l/32) of sample l%32, and the B operand of the next layer is taken
// from those elements directly: its K slot t of step q (lane half h) stands
// for unit U(h, 8q + t), and the weights (A operand, row l%32) are held in
// the same permuted order, in registers from the kernel start.  Inputs and
// activations enter each product as bf16, accumulation is fp32; layer 1 and
// layer 3 start from their biases, layer 2 adds its bias after the sum --
// the rounding points of the training kernel.  The three softmax weights
// overwrite the row's first three inputs.
typedef __bf16 mlpBf16x8 __attribute__((ext_vector_type(8)));
typedef short mlpI16x8 __attribute__((ext_vector_type(8)));
typedef float mlpF32x16 __attribute__((ext_vector_type(16)));

__device__ inline int mlpUnit(int h, int r) { return 8 * (r >> 2) + 4 * h + (r & 3); }
// fp32 -> bf16 bits, round to nearest even: the hardware conversion
// (v_cvt_pk_bf16_f32), the same value as bf16Bits for finite inputs
__device__ inline short mlpBf16(float x) { return __builtin_bit_cast(short, (__bf16)x); }

// LDS written by one lane of a wave and read by another: the wave's own
// accesses in order
__device__ inline void waveSyncLds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// This lane's operands: the three weight matrices in the permuted order
// and the biases of its accumulator elements.
struct MlpLane {
  mlpBf16x8 w1, w2[2], w3[2];
  float b1[16], b2[16], b3[16];
};
__device__ __attribute__((always_inline)) inline MlpLane mlpLane(const uint16_t *P) {
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  MlpLane o;
  mlpI16x8 a;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int i = 8 * h + t;
    a[t] = (short)((i < MLP_IN) ? P[MLP_W1 + col * MLP_IN + i] : 0);
  }
  o.w1 = __builtin_bit_cast(mlpBf16x8, a);
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    mlpI16x8 b, c;
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int u = mlpUnit(h, 8 * q + t);
      b[t] = (short)P[MLP_W2 + col * MLP_H + u];
      c[t] = (short)((col < MLP_OUT) ? P[MLP_W3 + col * MLP_H + u] : 0);
    }
    o.w2[q] = __builtin_bit_cast(mlpBf16x8, b);
    o.w3[q] = __builtin_bit_cast(mlpBf16x8, c);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int u = mlpUnit(h, r);
    o.b1[r] = bf16Float(P[MLP_B1 + u]);
    o.b2[r] = bf16Float(P[MLP_B2 + u]);
    o.b3[r] = (u < MLP_OUT) ? bf16Float(P[MLP_B3 + u]) : 0.f;
  }
  return o;
}

// Two blocks of 32 rows at a time, so that the matrix and activation work
// of one overlaps the other's latency.
template <int NROWS>
__device__ __attribute__((always_inline)) void waveMlp(float *rows, const MlpLane &W) {
  static_assert(NROWS % 64 == 0, "pairs of 32-row blocks");
  const int lane = threadIdx.x & 63, col = lane & 31, h = lane >> 5;
  for (int rb = 0; rb < NROWS / 32; rb += 2) {
    float *row[2] = {rows + (rb * 32 + col) * MLP_IN, rows + ((rb + 1) * 32 + col) * MLP_IN};
    // rows without the network hold zeros: a pair of blocks without any
    // network row is skipped
    bool any = false;
#pragma unroll
    for (int t = 0; t < MLP_IN; ++t) any = any || row[0][t] != 0.f || row[1][t] != 0.f;
    if (!__any(any)) continue;
    mlpF32x16 acc[2];
    mlpI16x8 g0[2], g1[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      mlpI16x8 xb;
#pragma unroll
      for (int t = 0; t < 8; ++t) xb[t] = (h == 0 && t < MLP_IN) ? mlpBf16(row[p][t]) : (short)0;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[p][r] = W.b1[r];
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w1, __builtin_bit_cast(mlpBf16x8, xb), acc[p], 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        g0[p][t] = mlpBf16(mlpGelu(acc[p][t]));
        g1[p][t] = mlpBf16(mlpGelu(acc[p][8 + t]));
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[p][r] = 0.f;
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w2[0], __builtin_bit_cast(mlpBf16x8, g0[p]), acc[p], 0, 0, 0);
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w2[1], __builtin_bit_cast(mlpBf16x8, g1[p]), acc[p], 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        g0[p][t] = mlpBf16(mlpGelu(acc[p][t] + W.b2[t]));
        g1[p][t] = mlpBf16(mlpGelu(acc[p][8 + t] + W.b2[8 + t]));
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[p][r] = W.b3[r];
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w3[0], __builtin_bit_cast(mlpBf16x8, g0[p]), acc[p], 0, 0, 0);
      acc[p] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W.w3[1], __builtin_bit_cast(mlpBf16x8, g1[p]), acc[p], 0, 0, 0);
    }
    if (h == 0) {  // logits of sample col: units 0, 1, 2 = elements 0, 1, 2
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const float m = fmaxf(acc[p][0], fmaxf(acc[p][1], acc[p][2]));
        const float e0 = mlpExp(acc[p][0] - m), e1 = mlpExp(acc[p][1] - m), e2 = mlpExp(acc[p][2] - m);
        const float sum = (e0 + e1) + e2;
        row[p][0] = e0 / sum;
        row[p][1] = e1 / sum;
        row[p][2] = e2 / sum;
      }
    }
  }
  waveSyncLds();
}

template <int D, int SCHEME, bool SEG, bool LWAVE = false>  // LWAVE: learned weno5, network per wave
__global__ __launch_bounds__(256) void kRates(BandView b, const uint32_t *list, const uint32_t *nact, RateParams rp,
                                              double *rates, DevScalars *sc, SegArgs sa, unsigned long long *segMax) {
  const uint32_t n = *nact;
  LSG_ABORT_IF_OVERFLOW(sc);
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  const Vel vel = velOf(rp);
  // learned weno5 with the network per lane: the bf16 parameters are
  // widened to fp32 once per block
  __shared__ float mlpW[(SCHEME == S_WENO5L && !LWAVE) ? MLP_NPARAM : 1];
  if (SCHEME == S_WENO5L && !LWAVE) {
    for (int i = threadIdx.x; i < MLP_NPARAM; i += TPB) mlpW[i] = bf16Float(rp.mlp[i]);
    __syncthreads();
  }
  const uint32_t segFirst = (SEG && lo < hi) ? sa.iseg[list[lo]] : 0u;
  __shared__ float mlpIO[SCHEME == S_WENO5L ? TPB * 2 * D * MLP_IN : 1];
  float *io = (SCHEME == S_WENO5L) ? &mlpIO[threadIdx.x * 2 * D * MLP_IN] : nullptr;
  RatePointFn<D, SCHEME, SEG> fn{b,   rp,     vel,      rates, mlpW, io, {0, 0, 0}, {0, 0, 0}, 0., 0ull,
                                 sa,  segMax, segFirst, 0ull,  0,    nullptr};
  if (SCHEME == S_LF)
    for (int i = 0; i < 3; ++i)
      fn.ga[i] = __longlong_as_double((long long)(SEG ? rp.segGA[3 * segFirst + i] : sc->alphaBits[i]));
  if constexpr (SEG) {
    for (uint32_t j = lo + threadIdx.x; j < hi; j += TPB) fn(list[j], j);
  } else {
    // Each XCD takes a contiguous slab of the active list in proportion to
    // its blocks (the stencil reads of neighbouring points then meet in that
    // XCD's L2), dealt in 64-point chunks to the XCD's waves in launch order
    // so that few active points still fill whole waves on every CU.
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t xcd = blockIdx.x % NUM_XCD, k = blockIdx.x / NUM_XCD;
    const uint32_t per = gridDim.x / NUM_XCD, rem = gridDim.x % NUM_XCD;
    const uint32_t nbx = per + (xcd < rem ? 1u : 0u), b0 = xcd * per + (xcd < rem ? xcd : rem);
    const uint32_t first = (uint32_t)((uint64_t)n * b0 / gridDim.x);
    const uint32_t last = (uint32_t)((uint64_t)n * (b0 + nbx) / gridDim.x);
    const uint32_t wave = k * (TPB / 64) + (threadIdx.x >> 6), waves = nbx * (TPB / 64);
    if constexpr (SCHEME == S_WENO5L && LWAVE) {
      // every lane of the wave takes part in the network of each 64 points
      const MlpLane W = mlpLane(rp.mlp);
      float *rows = &mlpIO[(threadIdx.x & ~63u) * 2 * D * MLP_IN];
      __shared__ double stCache[LWAVE ? TPB * D * 7 : 1];  // lane-interleaved per wave
      fn.stc = &stCache[(threadIdx.x & ~63u) * D * 7 + lane];
      fn.learnedPhase = 2;
      for (uint32_t base = first + wave * 64u; base < last; base += waves * 64u) {
        const uint32_t j = base + lane;
        if (j < last) fn.prepareLearned(list[j]);
        else
          for (int q = 0; q < 2 * D * MLP_IN; ++q) io[q] = 0.f;
        waveSyncLds();
        waveMlp<64 * 2 * D>(rows, W);
        if (j < last) fn(list[j], j);
        waveSyncLds();  // the rows are rewritten by the next points
      }
    } else {
      for (uint32_t base = first + wave * 64u; base < last; base += waves * 64u) {
        const uint32_t j = base + lane;
        if (j < last) fn(list[j], j);
      }
    }
  }
  if (!SEG) {
    blockMaxAtomic((unsigned long long)__double_as_longlong(fn.lmax), &sc->maxAbsBits);
    if (SCHEME == S_LLLF || SCHEME == S_LLF || SCHEME == S_LLFA)
      for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(fn.la[d]), &sc->alphaBits[d]);
    if (SCHEME == S_SLLF) blockMaxAtomic(fn.sllfKey, &sc->sllfKey);
  } else {
    // the level sets of a block's chunk are few and consecutive: reduce the
    // lanes' maxima in LDS slots, then one atomic per level set
    constexpr int NSLOT = 8;
    __shared__ unsigned long long slot[NSLOT];
    if (threadIdx.x < NSLOT) slot[threadIdx.x] = 0ull;
    __syncthreads();
    if (fn.lmax > 0.) {
      const unsigned long long bits = (unsigned long long)__double_as_longlong(fn.lmax);
      const uint32_t d = fn.curSeg - segFirst;
      if (d < NSLOT) atomicMax(&slot[d], bits);
      else atomicMax(&segMax[fn.curSeg], bits);
    }
    __syncthreads();
    if (threadIdx.x < NSLOT && slot[threadIdx.x]) atomicMax(&segMax[segFirst + threadIdx.x], slot[threadIdx.x]);
    if (SCHEME == S_LLLF || SCHEME == S_LLF)
      for (int d = 0; d < D; ++d)
        if (fn.la[d] > 0.)
          atomicMax(&rp.segAlpha[3 * fn.curSeg + d], (unsigned long long)__double_as_longlong(fn.la[d]));
  }
  unsigned long long t = blockSumU64(fn.act);
  if (threadIdx.x == 0 && t) atomicAdd(&sc->activeCount, t);
}

// StencilLocalLaxFriedrichsScalar::reduceTimeStepHamiltonJacobi
// (lsStencilLocalLaxFriedrichsScalar.hpp:385-397) uses the finalAlphas of the
// last point (entry order) whose scheme call set them: that call is evaluated
// again here (same inputs, same result) and its alphas become the stage's.
template <int D>
__global__ void kSllfAlphas(BandView b, const uint32_t *list, RateParams rp, DevScalars *sc, int severalLevelSets) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const unsigned long long key = sc->sllfKey;
  double fa[3] = {0., 0., 0.};
  if (key) {
    const uint32_t j = (uint32_t)(key >> 4) - 1u;
    const int mat = (int)(key & 15u);
    const uint32_t i = list[j];
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    const Vel base = velOf(rp);
    const Vel vel = severalLevelSets ? velOfMaterial(rp, base, mat) : base;
    bool used = false;
    (void)schemeSLLF<D>(b, c, vel, rp.sllfBaseEps, fa, used);
  }
  for (int d = 0; d < 3; ++d) sc->alphaBits[d] = (unsigned long long)__double_as_longlong(fa[d]);
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
  sc->sllfKey = 0;
  if (!keepAlphas) sc->alphaBits[0] = sc->alphaBits[1] = sc->alphaBits[2] = 0;
  sc->activeCount = 0;
}

// Active list of band b into ws.cpos / ws.ncand (both free between band operations).
static void activeList(hipStream_t s, const BandBufs &b, Workspace &ws, DevScalars *sc, int resetRates = 0,
                       int keepAlphas = 0) {
  kActiveCount<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials, sc, resetRates, keepAlphas);
  kActiveList<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials, ws.cpos, ws.ncand, sc);
}

void launchRates(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const RateParams &rp,
                 double *rates, DevScalars *sc, const double *maxTimeStep, double *dtOut, Workspace &ws,
                 hipEvent_t e0, hipEvent_t e1, bool externalEvents) {
  const int scheme = schemeIdOf(rp.spatial);
  if (!rp.hostVel) activeList(s, b, ws, sc, 1, scheme == S_LF ? 1 : 0);  // resets the rate maxima too
  else kResetRates<<<1, 1, 0, s>>>(sc, scheme == S_LF ? 1 : 0);  // the list of launchVelInputs is still in ws
  BandView v = viewOf(b, f, g);
  const int ge = rateBlocks();
  recordEvent(s, e0, externalEvents);
#define LSG_RATES(DD, SS) \
  kRates<DD, SS, false><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, rp, rates, sc, noSeg(), nullptr)
  if (D == 2) {
    switch (scheme) {
    case S_EO: LSG_RATES(2, S_EO); break;
    case S_WENO5: LSG_RATES(2, S_WENO5); break;
    case S_WENO5L:
      if (rp.learnedWave)
        kRates<2, S_WENO5L, false, true><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, rp, rates, sc, noSeg(), nullptr);
      else
        LSG_RATES(2, S_WENO5L);
      break;
    case S_LF: LSG_RATES(2, S_LF); break;
    case S_LLLF: LSG_RATES(2, S_LLLF); break;
    case S_LLFA: LSG_RATES(2, S_LLFA); break;
    case S_LLF: LSG_RATES(2, S_LLF); break;
    case S_SLLF: LSG_RATES(2, S_SLLF); break;
    default: LSG_RATES(2, S_WENO3); break;
    }
  } else {
    switch (scheme) {
    case S_EO: LSG_RATES(3, S_EO); break;
    case S_WENO5: LSG_RATES(3, S_WENO5); break;
    case S_WENO5L:
      if (rp.learnedWave)
        kRates<3, S_WENO5L, false, true><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, rp, rates, sc, noSeg(), nullptr);
      else
        LSG_RATES(3, S_WENO5L);
      break;
    case S_LF: LSG_RATES(3, S_LF); break;
    case S_LLLF: LSG_RATES(3, S_LLLF); break;
    case S_LLFA: LSG_RATES(3, S_LLFA); break;
    case S_LLF: LSG_RATES(3, S_LLF); break;
    case S_SLLF: LSG_RATES(3, S_SLLF); break;
    default: LSG_RATES(3, S_WENO3); break;
    }
  }
#undef LSG_RATES
  recordEvent(s, e1, externalEvents);
  if (scheme == S_SLLF) {
    if (D == 2) kSllfAlphas<2><<<1, 1, 0, s>>>(v, ws.cpos, rp, sc, 0);
    else kSllfAlphas<3><<<1, 1, 0, s>>>(v, ws.cpos, rp, sc, 0);
  }
  const int useAlphas =
      (scheme == S_LF || scheme == S_LLLF || scheme == S_LLF || scheme == S_LLFA || scheme == S_SLLF) ? 1 : 0;
  kTimeStep<<<1, 1, 0, s>>>(sc, rp, D, g.dx, maxTimeStep, useAlphas, dtOut, b.n);
}

// ------------------------------------------------ several level sets ----

// integrateTime with level sets below the top (lsAdvect.hpp:496-593): for
// each active point of the top, the material is the first level set
// (bottom up, the top included) whose value there is <= value + 1e-4; an
// etching point closer to the level set below than the remaining CFL
// budget stops on it and continues with the material there.  Writes the
// point's (velocity after the dissipation rule, stop value) list and
// reduces the minimum step limit.
template <int D, int SCHEME>
__global__ __launch_bounds__(256) void kRatesMM(BandView b, LowerViews lw, const uint32_t *list, const uint32_t *nact,
                                                RateParams rp, double *lists, DevScalars *sc) {
  const uint32_t n = *nact;
  LSG_ABORT_IF_OVERFLOW(sc);
  const Vel base = velOf(rp);
  const int L = lw.n + 1;
  const double MAXV = 1.7976931348623157e308;
  double ga[3] = {0, 0, 0}, la[3] = {0, 0, 0};
  if (SCHEME == S_LF)
    for (int i = 0; i < 3; ++i) ga[i] = __longlong_as_double((long long)sc->alphaBits[i]);
  double minStep = MAXV;
  unsigned long long act = 0;
  unsigned long long sllfKey = 0;  // SLLF: (j + 1) << 4 | material of this thread's last call that set alphas
  for (uint32_t j = blockIdx.x * TPB + threadIdx.x; j < n; j += gridDim.x * TPB) {
    const uint32_t i = list[j];
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    const double phi0 = b.val[i];
    ++act;
    // values of every level set at c (the lower ones share the top's frame)
    double lv[MAX_LEVEL_SETS];
    const BandView::Loc l = b.locate(c[0], c[1], c[2]);
    LSG_UNROLL
    for (int m = 0; m < MAX_LEVEL_SETS - 1; ++m) {
      if (m < lw.n) {
        BandView v = b;
        v.bits = lw.bits[m];
        v.pre = lw.pre[m];
        v.val = lw.val[m];
        lv[m] = v.val[v.entryOf(l, v.bits[l.w], v.pre[l.w])];
      }
    }
    lv[L - 1] = phi0;
    const bool isVoid = rp.voidMarks && rp.voidMarks[rp.voidRank[i]] != 0.;
    double value = phi0, cfl = rp.cfl, maxStepTime = 0.;
    double *out = lists + (size_t)i * 2 * L;
    int k = 0;
    for (int cur = L - 1; cur >= 0; --cur) {  // :507-589
      int mat = -1;
      for (int m = 0; m < L; ++m)
        if (lv[m] <= value + 1e-4) {  // wrappingLayerEpsilon (:80)
          mat = m;
          break;
        }
      double velocity = 0., v = 0., dn = 0.;
      if (mat >= 0 && !isVoid) {
        Vel vm = velOfMaterial(rp, base, mat);
        if (rp.hostVel) {  // evaluated on the host for (active point j, material mat)
          const double *h = rp.hostVel + ((size_t)j * rp.hostRec * L + mat) * 4;
          vm.kind = 0;
          vm.b = h[0];
          vm.vec[0] = h[1];
          vm.vec[1] = h[2];
          vm.vec[2] = h[3];
          if (SCHEME == S_LLF || SCHEME == S_LLFA || SCHEME == S_SLLF) {
            vm.box = h + 4 * L;
            vm.boxStride = 4 * L;
          }
          if (SCHEME == S_SLLF) vm.sllf = rp.hostSllf + ((size_t)j * L + mat) * ((D == 3) ? 27 : 9) * 2 * D;
        }
        bool used = false;
        v = ratePoint<D, SCHEME>(b, c, phi0, rp, vm, ga, la, velocity, nullptr, nullptr, &used, &dn);
        if (SCHEME == S_SLLF && used) sllfKey = ((unsigned long long)(j + 1) << 4) | (unsigned long long)mat;
      }
      double stop;
      bool done = true;
      if (velocity > 0.) {
        maxStepTime += cfl / velocity;
        stop = -MAXV;
      } else if (velocity == 0.) {
        maxStepTime = MAXV;
        stop = MAXV;
      } else {
        const double valueBelow = (cur > 0) ? lv[cur - 1] : MAXV;
        const double difference = fabs(valueBelow - value);
        if (difference >= cfl) {
          maxStepTime -= cfl / velocity;
          stop = MAXV;
        } else if (rp.adaptive && difference > rp.adaptiveFactor * cfl) {  // sub-steps near the layer (:568-578)
          maxStepTime -= rp.adaptiveFactor * cfl / velocity;
          stop = MAXV;
        } else {  // terminal step onto the material below
          cfl -= difference;
          value = valueBelow;
          maxStepTime -= difference / velocity;
          stop = valueBelow;
          done = false;
        }
      }
      out[2 * k] = v;
      out[2 * k + 1] = stop;
      if (rp.dissOut) rp.dissOut[(size_t)i * L + k] = dn;
      ++k;
      if (done) break;
    }
    if (maxStepTime < minStep) minStep = maxStepTime;
  }
  // step limits are >= 0, so their bit patterns order like the values
  {
    unsigned long long mb = (unsigned long long)__double_as_longlong(minStep);
    __shared__ unsigned long long sm[TPB / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long y = __shfl_down(mb, o, 64);
      mb = (y < mb) ? y : mb;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sm[wid] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = sm[0];
      for (int q = 1; q < TPB / 64; ++q) m = (sm[q] < m) ? sm[q] : m;
      atomicMin(&sc->minStepBits, m);
    }
    __syncthreads();
  }
  if (SCHEME == S_LLLF || SCHEME == S_LLF || SCHEME == S_LLFA)
    for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(la[d]), &sc->alphaBits[d]);
  if (SCHEME == S_SLLF) blockMaxAtomic(sllfKey, &sc->sllfKey);
  unsigned long long t = blockSumU64(act);
  if (threadIdx.x == 0 && t) atomicAdd(&sc->activeCount, t);
}

__global__ void kResetMM(DevScalars *sc) { sc->minStepBits = (unsigned long long)__double_as_longlong(1.7976931348623157e308); }

__global__ void kTimeStepMM(DevScalars *sc, int D, double dx, const double *maxTimeStepPtr, int useAlphas,
                            double *dtOut) {
  LSG_ABORT_IF_OVERFLOW(sc);
  double al[3];
  for (int i = 0; i < 3; ++i) al[i] = __longlong_as_double((long long)sc->alphaBits[i]);
  const double m = __longlong_as_double((long long)sc->minStepBits);
  *dtOut = stageTimeStepMM(*maxTimeStepPtr, m, D, al, useAlphas != 0, dx);
  sc->totalActive += sc->activeCount;
}

void launchRatesMM(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const LowerViews &lw,
                   const RateParams &rp, double *lists, DevScalars *sc, const double *maxTimeStep, double *dtOut,
                   Workspace &ws) {
  const int scheme = schemeIdOf(rp.spatial);
  activeList(s, b, ws, sc, 1, scheme == S_LF ? 1 : 0);  // resets the rate maxima (not LF's global alphas)
  kResetMM<<<1, 1, 0, s>>>(sc);
  BandView v = viewOf(b, f, g);
  const int ge = rateBlocks();
#define LSG_RATES_MM(DD, SS) kRatesMM<DD, SS><<<ge, TPB, 0, s>>>(v, lw, ws.cpos, ws.ncand, rp, lists, sc)
  if (D == 2) {
    switch (scheme) {
    case S_EO: LSG_RATES_MM(2, S_EO); break;
    case S_LF: LSG_RATES_MM(2, S_LF); break;
    case S_WENO5: LSG_RATES_MM(2, S_WENO5); break;
    case S_LLLF: LSG_RATES_MM(2, S_LLLF); break;
    case S_LLFA: LSG_RATES_MM(2, S_LLFA); break;
    case S_LLF: LSG_RATES_MM(2, S_LLF); break;
    case S_SLLF: LSG_RATES_MM(2, S_SLLF); break;
    default: LSG_RATES_MM(2, S_WENO3); break;
    }
  } else {
    switch (scheme) {
    case S_EO: LSG_RATES_MM(3, S_EO); break;
    case S_LF: LSG_RATES_MM(3, S_LF); break;
    case S_WENO5: LSG_RATES_MM(3, S_WENO5); break;
    case S_LLLF: LSG_RATES_MM(3, S_LLLF); break;
    case S_LLFA: LSG_RATES_MM(3, S_LLFA); break;
    case S_LLF: LSG_RATES_MM(3, S_LLF); break;
    case S_SLLF: LSG_RATES_MM(3, S_SLLF); break;
    default: LSG_RATES_MM(3, S_WENO3); break;
    }
  }
#undef LSG_RATES_MM
  if (scheme == S_SLLF) {
    if (D == 2) kSllfAlphas<2><<<1, 1, 0, s>>>(v, ws.cpos, rp, sc, 1);
    else kSllfAlphas<3><<<1, 1, 0, s>>>(v, ws.cpos, rp, sc, 1);
  }
  const int useAlphas =
      (scheme == S_LF || scheme == S_LLLF || scheme == S_LLF || scheme == S_LLFA || scheme == S_SLLF) ? 1 : 0;
  kTimeStepMM<<<1, 1, 0, s>>>(sc, D, g.dx, maxTimeStep, useAlphas, dtOut);
}

__global__ void kSetFlag(uint32_t *flag, uint32_t v) { *flag = v; }
void launchSetFlag(hipStream_t s, uint32_t *flag, uint32_t v) { kSetFlag<<<1, 1, 0, s>>>(flag, v); }

// Active list, defined list and the velocity inputs of the active points.
void launchVelInputs(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const RateParams &rp,
                     const LowerViews &lw, Workspace &ws, uint32_t *dlist, uint32_t *ndef, double *out,
                     DevScalars *sc) {
  activeList(s, b, ws, sc);
  kDefinedCount<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials);
  kDefinedList<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials, dlist, ndef);
  BandView v = viewOf(b, f, g);
  if (D == 2) kVelInputs<2><<<rateBlocks(), TPB, 0, s>>>(v, lw, ws.cpos, ws.ncand, dlist, ndef, rp, out);
  else kVelInputs<3><<<rateBlocks(), TPB, 0, s>>>(v, lw, ws.cpos, ws.ncand, dlist, ndef, rp, out);
}

// Batch rate pass: every level set's stage velocities and max |H - D| per
// level set; then the per-level-set time steps (EO, WENO and the
// Lax-Friedrichs schemes LF, LLF and LLLF).
__global__ void kTimeStepSeg(DevScalars *sc, const SegInfo *segs, uint32_t nseg, double cfl, int D,
                             const unsigned long long *segMax, const double *maxDt, double *dtOut, const uint32_t *nb,
                             const unsigned long long *alphas) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nseg) {
    const double m = __longlong_as_double((long long)segMax[k]);
    double al[3] = {0., 0., 0.};
    if (alphas)
      for (int d = 0; d < 3; ++d) al[d] = __longlong_as_double((long long)alphas[3 * k + d]);
    dtOut[k] = stageTimeStep(maxDt[k], cfl, m, D, al, alphas != nullptr, segs[k].g.dx);
  }
  if (k == 0) {
    sc->totalActive += sc->activeCount;
    sc->rateBytes += 8.0 * (double)(*nb) + 8.0 * (double)sc->activeCount;
  }
}
__global__ void kResetRatesSeg(DevScalars *sc, unsigned long long *segMax, uint32_t nseg, unsigned long long *segAlpha,
                               unsigned long long *segGA) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nseg) {
    segMax[k] = 0ull;
    for (int d = 0; d < 3; ++d) {
      if (segAlpha) segAlpha[3 * k + d] = 0ull;
      if (segGA) segGA[3 * k + d] = 0ull;
    }
  }
  if (k == 0) sc->activeCount = 0;
}

// Advect::findGlobalAlphas (lsAdvect.hpp:90-195) per level set of a batch:
// the maxima over each level set's active points with its own functor.
// A lane's running maxima belong to its current level set and are flushed
// when that changes (the active list holds each level set's points
// together).
template <int D>
__global__ __launch_bounds__(256) void kGlobalAlphasSeg(BandView b, const uint32_t *list, const uint32_t *nact,
                                                        SegArgs sa, unsigned long long *segGA, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nact;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t cur = 0xFFFFFFFFu;
  double la[3] = {0., 0., 0.};
  auto flush = [&]() {
    if (cur == 0xFFFFFFFFu) return;
    for (int d = 0; d < D; ++d)
      if (la[d] > 0.) atomicMax(&segGA[3 * cur + d], (unsigned long long)__double_as_longlong(la[d]));
  };
  for (uint32_t j = lo + threadIdx.x; j < hi; j += TPB) {
    const uint32_t i = list[j];
    const uint32_t k = sa.iseg[i];
    if (k != cur) {
      flush();
      la[0] = la[1] = la[2] = 0.;
      cur = k;
    }
    const SegInfo &S = sa.segs[k];
    const BandView v = segView(b, S);
    int32_t c[3];
    candCoords(v, i, v.pos[i], c);
    globalAlphaPoint<D>(v, c, S.vel, la);
  }
  flush();
}

void launchRatesSeg(hipStream_t s, int D, const BandBufs &b, const SegCtx &c, const RateParams &rp, double *rates,
                    DevScalars *sc, unsigned long long *segMax, const double *maxDt, double *dtOut, Workspace &ws) {
  const int scheme = schemeIdOf(rp.spatial);
  const int gs = (int)((c.nseg + TPB - 1) / TPB);
  kResetRatesSeg<<<gs, TPB, 0, s>>>(sc, segMax, c.nseg, rp.segAlpha, rp.segGA);
  activeList(s, b, ws, sc);
  Frame f{};
  GridInfo g{};
  BandView v = viewOf(b, f, g);
  SegArgs sa = segArgs(c, b.seg, nullptr);
  const int ge = rateBlocks();
  if (scheme == S_LF) {
    if (D == 2) kGlobalAlphasSeg<2><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, sa, rp.segGA, sc);
    else kGlobalAlphasSeg<3><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, sa, rp.segGA, sc);
  }
#define LSG_RATES_SEG(DD, SS) kRates<DD, SS, true><<<ge, TPB, 0, s>>>(v, ws.cpos, ws.ncand, rp, rates, sc, sa, segMax)
  if (D == 2) {
    switch (scheme) {
    case S_EO: LSG_RATES_SEG(2, S_EO); break;
    case S_WENO3: LSG_RATES_SEG(2, S_WENO3); break;
    case S_WENO5L: LSG_RATES_SEG(2, S_WENO5L); break;
    case S_LF: LSG_RATES_SEG(2, S_LF); break;
    case S_LLLF: LSG_RATES_SEG(2, S_LLLF); break;
    case S_LLF: LSG_RATES_SEG(2, S_LLF); break;
    default: LSG_RATES_SEG(2, S_WENO5); break;
    }
  } else {
    switch (scheme) {
    case S_EO: LSG_RATES_SEG(3, S_EO); break;
    case S_WENO3: LSG_RATES_SEG(3, S_WENO3); break;
    case S_WENO5L: LSG_RATES_SEG(3, S_WENO5L); break;
    case S_LF: LSG_RATES_SEG(3, S_LF); break;
    case S_LLLF: LSG_RATES_SEG(3, S_LLLF); break;
    case S_LLF: LSG_RATES_SEG(3, S_LLF); break;
    default: LSG_RATES_SEG(3, S_WENO5); break;
    }
  }
#undef LSG_RATES_SEG
  // the per-level-set step limits of the Lax-Friedrichs schemes use their
  // alphas (global ones for LF, the pass's local maxima for LLF and LLLF)
  const unsigned long long *alphas =
      (scheme == S_LF) ? rp.segGA : ((scheme == S_LLLF || scheme == S_LLF) ? rp.segAlpha : nullptr);
  kTimeStepSeg<<<gs, TPB, 0, s>>>(sc, c.segs, c.nseg, rp.cfl, D, segMax, maxDt, dtOut, b.n, alphas);
}

// Per-level-set step bookkeeping of a batch: stage-1 bound, returned time of
// the step (lsAdvectIntegrationSchemes.hpp:58,107,167) and advected time.
__global__ void kStepBeginSeg(DevScalars *sc, double *maxDt, uint32_t nseg) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nseg) maxDt[k] = 1.7976931348623157e308;
  if (k == 0) {
    sc->overflow = 0;
    sc->stepActive0 = sc->totalActive;
    sc->reframe = 0;
  }
}
__global__ void kStepTimeSeg(const DevScalars *sc, const double *dt, uint32_t nseg, int temporal, double *stepTime,
                             double *advTime) {
  LSG_ABORT_IF_OVERFLOW(sc);
  if (sc->lookbackFail) return;  // the step is repeated
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nseg) return;
  const double d1 = dt[k], d2 = dt[nseg + k], d3 = dt[2 * nseg + k];
  double t;
  if (temporal == 1) t = 0.5 * d1 + 0.5 * d2;
  else if (temporal == 2) t = (d1 + d2 + 4.0 * d3) / 6.0;
  else t = d1;
  stepTime[k] = t;
  advTime[k] += t;
}
void launchStepBeginSeg(hipStream_t s, DevScalars *sc, double *maxDt, uint32_t nseg) {
  kStepBeginSeg<<<(nseg + TPB - 1) / TPB, TPB, 0, s>>>(sc, maxDt, nseg);
}
void launchStepTimeSeg(hipStream_t s, const DevScalars *sc, const double *dt, uint32_t nseg, int temporal,
                       double *stepTime, double *advTime) {
  kStepTimeSeg<<<(nseg + TPB - 1) / TPB, TPB, 0, s>>>(sc, dt, nseg, temporal, stepTime, advTime);
}

// Flags (sc->reframe = 1) an entry closer than `slack` to the frame edge in
// an infinite dimension (the host then moves the frame between steps).
__global__ void kFrameCheck(const uint32_t *pos, const uint32_t *nptr, Frame f, GridInfo g, int D, int slack,
                            DevScalars *sc) {
  const uint32_t n = *nptr;
  bool near = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = pos[i];
    if (p == POS_NONE) continue;
    int32_t r[3];
    f.decode(p, r[0], r[1], r[2]);
    for (int d = 0; d < D; ++d)
      if (!g.bounded(d) && (r[d] < slack || f.ext[d] - 1 - r[d] < slack)) near = true;
  }
  if (near) sc->reframe = 1;
}
void launchFrameCheck(hipStream_t s, const BandBufs &b, const Frame &f, const GridInfo &g, int D, int slack,
                      DevScalars *sc) {
  kFrameCheck<<<gridFor(b.cap, 2048), TPB, 0, s>>>(b.pos, b.n, f, g, D, slack, sc);
}

// Flags (sc->reframe = 1) an entry of a batch band closer than `slack` to its
// frame edge in an infinite dimension: the host then moves the frames.
__global__ void kFrameCheckSeg(const uint32_t *pos, const uint16_t *seg, const uint32_t *nptr, const SegInfo *segs,
                               int D, int slack, DevScalars *sc) {
  const uint32_t n = *nptr;
  bool near = false;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t p = pos[i];
    if (p == POS_NONE) continue;
    const SegInfo &S = segs[seg[i]];
    int32_t r[3];
    S.f.decode(p, r[0], r[1], r[2]);
    for (int d = 0; d < D; ++d)
      if (!S.g.bounded(d) && (r[d] < slack || S.f.ext[d] - 1 - r[d] < slack)) near = true;
  }
  if (near) sc->reframe = 1;
}
void launchFrameCheckSeg(hipStream_t s, const BandBufs &b, const SegCtx &c, int D, int slack, DevScalars *sc) {
  kFrameCheckSeg<<<gridFor(b.cap, 2048), TPB, 0, s>>>(b.pos, b.seg, b.n, c.segs, D, slack, sc);
}

// Advect::findGlobalAlphas (lsAdvect.hpp:90-195), one level set.
template <int D> struct AlphaPointFn {
  BandView b;
  Vel vel;
  double la[3];
  __device__ __attribute__((always_inline)) void operator()(uint32_t i) {
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    globalAlphaPoint<D>(b, c, vel, la);
  }
};

template <int D> __global__ __launch_bounds__(256) void kGlobalAlphas(BandView b, const uint32_t *nptr, RateParams rp, DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  const Vel vel = velOf(rp);
  AlphaPointFn<D> fn{b, vel, {0, 0, 0}};
  forEachActive(b.val, n, fn);
  for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(fn.la[d]), &sc->alphaBits[d]);
}

__global__ void kResetAlphas(DevScalars *sc) { sc->alphaBits[0] = sc->alphaBits[1] = sc->alphaBits[2] = 0; }

template <int D> struct AlphaPointMMFn {
  BandView b;
  LowerViews lw;
  RateParams rp;
  Vel base;
  double la[3];
  __device__ __attribute__((always_inline)) void operator()(uint32_t i) {
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    const double phi0 = b.val[i];
    double lv[MAX_LEVEL_SETS];
    levelValuesAt(b, lw, c, phi0, lv);
    const Vel vel = velOfMaterial(rp, base, alphaMaterial(lv, lw.n + 1, phi0));
    globalAlphaPoint<D>(b, c, vel, la);
  }
};
template <int D>
__global__ __launch_bounds__(256) void kGlobalAlphasMM(BandView b, LowerViews lw, const uint32_t *nptr, RateParams rp,
                                                       DevScalars *sc) {
  LSG_ABORT_IF_OVERFLOW(sc);
  const uint32_t n = *nptr;
  AlphaPointMMFn<D> fn{b, lw, rp, velOf(rp), {0, 0, 0}};
  forEachActive(b.val, n, fn);
  for (int d = 0; d < D; ++d) blockMaxAtomic((unsigned long long)__double_as_longlong(fn.la[d]), &sc->alphaBits[d]);
}
void launchGlobalAlphasMM(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g,
                          const LowerViews &lw, const RateParams &rp, DevScalars *sc) {
  kResetAlphas<<<1, 1, 0, s>>>(sc);
  BandView v = viewOf(b, f, g);
  const int ge = gridFor(b.cap, 8192);
  if (D == 2) kGlobalAlphasMM<2><<<ge, TPB, 0, s>>>(v, lw, b.n, rp, sc);
  else kGlobalAlphasMM<3><<<ge, TPB, 0, s>>>(v, lw, b.n, rp, sc);
}

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

__global__ void kStepBegin(DevScalars *sc, const double *maxDt) {
  sc->overflow = 0;
  sc->reframe = 0;
  sc->stepActive0 = sc->totalActive;
  sc->maxTimeStep = *maxDt;
}
void launchStepBegin(hipStream_t s, DevScalars *sc, const double *maxDt) { kStepBegin<<<1, 1, 0, s>>>(sc, maxDt); }

// ---------------------------------------------------------------- misc ----

__global__ void kCopyBand(const double *sv, const uint32_t *sp, const uint32_t *sn, double *dv, uint32_t *dp,
                          uint32_t *dn, size_t cap, const uint64_t *sb, const uint32_t *spre, uint64_t *db,
                          uint32_t *dpre, int64_t nw, const uint16_t *ss, uint16_t *ds) {
  uint32_t n = *sn;
  if (n > cap) n = (uint32_t)cap;
  size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = tid; i < n; i += stride) {
    dv[i] = sv[i];
    dp[i] = sp[i];
    if (ds) ds[i] = ss[i];
  }
  for (size_t w = tid; w <= (size_t)nw; w += stride) {
    db[w] = sb[w];
    dpre[w] = spre[w];
  }
  if (tid == 0) *dn = n;
}
void launchCopyBand(hipStream_t s, const BandBufs &src, BandBufs &dst, const Frame &f) { launchCopyBandWords(s, src, dst, f.nwords); }
void launchCopyBandWords(hipStream_t s, const BandBufs &src, BandBufs &dst, int64_t nw) {
  size_t work = src.cap > (size_t)nw ? src.cap : (size_t)nw;
  kCopyBand<<<gridFor(work, 8192), TPB, 0, s>>>(src.val, src.pos, src.n, dst.val, dst.pos, dst.n, dst.cap, src.bits,
                                                src.pre, dst.bits, dst.pre, nw, (src.seg && dst.seg) ? src.seg : nullptr,
                                                (src.seg && dst.seg) ? dst.seg : nullptr);
}

__device__ inline int32_t blockMinI(int32_t v) {
  __shared__ int32_t smn[TPB / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    int32_t y = __shfl_down(v, o, 64);
    v = y < v ? y : v;
  }
  if ((threadIdx.x & 63) == 0) smn[threadIdx.x >> 6] = v;
  __syncthreads();
  int32_t r = smn[0];
  for (int k = 1; k < TPB / 64; ++k) r = smn[k] < r ? smn[k] : r;
  __syncthreads();
  return r;
}

__global__ void kBandStats(const double *val, const uint32_t *pos, const uint32_t *nptr, Frame f, uint32_t *count,
                           int32_t *bbox, uint32_t first) {
  const uint32_t n = *nptr;
  int32_t mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  unsigned long long c = 0;
  for (uint32_t i = first + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    if (!isUndef(val[i])) ++c;
    if (pos[i] == POS_NONE) continue;
    int32_t r[3];
    f.decode(pos[i], r[0], r[1], r[2]);
    for (int d = 0; d < 3; ++d) {
      mn[d] = r[d] < mn[d] ? r[d] : mn[d];
      mx[d] = r[d] > mx[d] ? r[d] : mx[d];
    }
  }
  unsigned long long t = blockSumU64(c);
  int32_t bmn[3], bmx[3];
  for (int d = 0; d < 3; ++d) {
    bmn[d] = blockMinI(mn[d]);
    bmx[d] = -blockMinI(mx[d] == INT32_MIN ? INT32_MAX : -mx[d]);
  }
  if (threadIdx.x == 0) {
    if (t) atomicAdd(count, (uint32_t)t);
    for (int d = 0; d < 3; ++d) {
      if (bmn[d] != INT32_MAX) atomicMin(&bbox[d], bmn[d]);
      if (bmx[d] > INT32_MIN + 1) atomicMax(&bbox[3 + d], bmx[d]);
    }
  }
}
__global__ void kStatsInit(uint32_t *count, int32_t *bbox) {
  *count = 0;
  for (int d = 0; d < 3; ++d) {
    bbox[d] = INT32_MAX;
    bbox[3 + d] = INT32_MIN;
  }
}
void launchBandStats(hipStream_t s, const BandBufs &b, const Frame &f, uint32_t *outCount, int32_t *bbox,
                     uint32_t first) {
  kStatsInit<<<1, 1, 0, s>>>(outCount, bbox);
  kBandStats<<<gridFor(b.cap, 512), TPB, 0, s>>>(b.val, b.pos, b.n, f, outCount, bbox, first);
}

// BooleanOperation INVERT (lsBooleanOperation.hpp:214-247): every defined
// value negated and the two undefined run signs swapped, i.e. every entry of
// the run list negated; the run structure stays.
__global__ void kNegate(double *val, const uint32_t *nptr) {
  const uint32_t n = *nptr;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) val[i] = -val[i];
}
void launchNegate(hipStream_t s, const BandBufs &b) { kNegate<<<gridFor(b.cap), TPB, 0, s>>>(b.val, b.n); }

__global__ void kFill32(uint32_t *p, uint32_t v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
void launchFill32(hipStream_t s, uint32_t *p, uint32_t v, size_t n) { kFill32<<<gridFor(n), TPB, 0, s>>>(p, v, n); }

// ------------------------------------------------------------ point data ---
// PointData::translateFromMultiData after a band operation: the point id of
// every defined entry of the operation's input, then for every defined
// entry of the output (in order) the id it takes its data from (the
// newDataSourceIds of lsExpand.hpp:109-150, lsReduce.hpp:105-111,
// lsPrune.hpp:145-167, lsBooleanOperation.hpp:104-115, lsAdvect.hpp:304-413),
// then one gather per array.

__global__ void kRankScatter(const uint32_t *dlist, const uint32_t *ndef, uint32_t *rankOf) {
  const uint32_t n = *ndef;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) rankOf[dlist[k]] = k;
}
void launchPointIds(hipStream_t s, const BandBufs &b, Workspace &ws, uint32_t *dlist, uint32_t *ndef,
                    uint32_t *rankOf) {
  kDefinedCount<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials);
  kDefinedList<<<g_scanBlocks, TPB, 0, s>>>(b.val, b.n, ws.partials, dlist, ndef);
  kRankScatter<<<gridFor(b.cap), TPB, 0, s>>>(dlist, ndef, rankOf);
}

struct PdSrcArgs {
  BandView out, in, inB;  // the result, the operation's input (A), and B of a boolean
  const uint32_t *outN;   // entries of the result
  PdSrcParams p;
  const uint32_t *rankIn, *rankB;
};

// Source of defined output entry i: false when it records none.
template <int D>
__device__ inline bool pdSource(const PdSrcArgs &a, uint32_t i, uint32_t &id, uint8_t &side, uint32_t &entry) {
  if (isUndef(a.out.val[i])) return false;
  int32_t c[3];
  candCoords(a.out, i, a.out.pos[i], c);
  if (a.p.kind == SRC_BOOLEAN) {
    const BandView::Loc l = a.in.locate(c[0], c[1], c[2]);
    const uint32_t ea = a.in.entryOf(l, a.in.bits[l.w], a.in.pre[l.w]);
    const uint32_t eb = a.inB.entryOf(l, a.inB.bits[l.w], a.inB.pre[l.w]);  // same frame
    const bool fromA = boolFromA(a.p.boolKind, a.in.val[ea], a.inB.val[eb]);
    side = fromA ? 1 : 0;
    entry = fromA ? ea : eb;
    id = fromA ? a.rankIn[ea] : a.rankB[eb];
    return true;
  }
  int s = 2 * D;
  if (a.p.kind != SRC_CENTRE) {
    double cv, nb[2 * D];
    gatherStar<D>(a.in, c, cv, nb);
    if (a.p.kind == SRC_EXPAND) s = srcExpand<D>(cv, nb, a.p.totalLimit);
    else if (a.p.kind == SRC_REBUILD) s = srcRebuild<D>(cv, nb);
    else s = srcPrune<D>(cv, nb, a.p.kind == SRC_PRUNE_RSZ);
  }
  if (s < 0) return false;
  int32_t q[3] = {c[0], c[1], c[2]};
  if (s < 2 * D) {
    const int k = s % D;
    q[k] = a.in.g.map(k, (int64_t)c[k] + (s < D ? 1 : -1));
  }
  const BandView::Loc l = a.in.locate(q[0], q[1], q[2]);
  entry = a.in.entryOf(l, a.in.bits[l.w], a.in.pre[l.w]);
  id = a.rankIn[entry];
  side = 1;
  return true;
}

template <int D> __global__ __launch_bounds__(256) void kPdSrcCount(PdSrcArgs a, uint32_t *partials) {
  const uint32_t n = *a.outN;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t acc = 0;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += TPB) {
    uint32_t id, entry;
    uint8_t side;
    acc += pdSource<D>(a, i, id, side, entry) ? 1u : 0u;
  }
  const uint32_t t = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = t;
}
template <int D>
__global__ __launch_bounds__(256) void kPdSrcList(PdSrcArgs a, const uint32_t *partials, uint32_t *src, uint8_t *sideOut,
                                                  uint32_t *entryOut, uint32_t *nPush) {
  const uint32_t n = *a.outN;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t off = blockOffset(partials);
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t i = t + threadIdx.x;
    uint32_t id = 0, entry = 0;
    uint8_t side = 0;
    const bool push = (i < hi) && pdSource<D>(a, i, id, side, entry);
    uint32_t tot;
    uint64_t wm;
    const uint32_t q = off + blockBallotScan(push, tot, wm);
    if (push) {
      src[q] = id;
      sideOut[q] = side;
      entryOut[q] = entry;
    }
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) *nPush = off;
}

void launchPdSources(hipStream_t s, int D, const BandBufs &out, const BandBufs &in, const BandBufs *inB, const Frame &f,
                     const GridInfo &g, const PdSrcParams &p, Workspace &ws, const uint32_t *rankIn,
                     const uint32_t *rankB, uint32_t *src, uint8_t *side, uint32_t *entry, uint32_t *nPush) {
  PdSrcArgs a;
  a.out = viewOf(out, f, g);
  a.outN = out.n;
  a.in = viewOf(in, f, g);
  a.inB = inB ? viewOf(*inB, f, g) : a.in;
  a.p = p;
  a.rankIn = rankIn;
  a.rankB = rankB ? rankB : rankIn;
  if (D == 2) {
    kPdSrcCount<2><<<g_scanBlocks, TPB, 0, s>>>(a, ws.partials);
    kPdSrcList<2><<<g_scanBlocks, TPB, 0, s>>>(a, ws.partials, src, side, entry, nPush);
  } else {
    kPdSrcCount<3><<<g_scanBlocks, TPB, 0, s>>>(a, ws.partials);
    kPdSrcList<3><<<g_scanBlocks, TPB, 0, s>>>(a, ws.partials, src, side, entry, nPush);
  }
}

// out[k] = the source's tuple (A where side[k], else B; 0 without B or past
// the end of the source array)
__global__ void kPdGather(int comps, const double *A, const uint32_t *lenA, const double *B, const uint32_t *lenB,
                          const uint32_t *src, const uint8_t *side, const uint32_t *nPush, double *out) {
  const uint32_t n = *nPush, la = *lenA, lb = lenB ? *lenB : 0u;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t id = src[k];
    const double *from = side[k] ? A : B;
    const bool ok = from && id < (side[k] ? la : lb);
    for (int j = 0; j < comps; ++j) out[(size_t)k * comps + j] = ok ? from[(size_t)id * comps + j] : 0.0;
  }
}
void launchPdGather(hipStream_t s, int comps, const double *A, const uint32_t *lenA, const double *B,
                    const uint32_t *lenB, const uint32_t *src, const uint8_t *side, const uint32_t *nPush, size_t cap,
                    double *out) {
  kPdGather<<<gridFor(cap), TPB, 0, s>>>(comps, A, lenA, B, lenB, src, side, nPush, out);
}

__global__ void kCopyU32(const uint32_t *from, uint32_t *to) { *to = *from; }
void launchCopyU32(hipStream_t s, const uint32_t *from, uint32_t *to) { kCopyU32<<<1, 1, 0, s>>>(from, to); }

// saveAdvectionVelocities (lsAdvect.hpp:733-796): per point of the reduced
// band (slot k, input entry entry[k]) the last rate = time * velocity and the
// dissipation of that list entry; one level set: rate = dt * rates[e].
__global__ void kPdVelocities(const uint32_t *entry, const uint32_t *nPush, const double *val, const double *rates,
                              const double *diss, const double *dtp, int nmat, double *vel, double *dis) {
  const uint32_t n = *nPush;
  const double dt = *dtp;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t e = entry[k];
    if (nmat <= 1) {
      vel[k] = dt * rates[e];
      dis[k] = diss[e];
    } else {
      int last = 0;
      vel[k] = updateRateMM(val[e], dt, rates + (size_t)e * 2 * nmat, last);
      dis[k] = diss[(size_t)e * nmat + last];
    }
  }
}
void launchPdVelocities(hipStream_t s, const uint32_t *entry, const uint32_t *nPush, const double *val,
                        const double *rates, const double *diss, const double *dt, int nmat, size_t cap, double *vel,
                        double *dis) {
  kPdVelocities<<<gridFor(cap), TPB, 0, s>>>(entry, nPush, val, rates, diss, dt, nmat, vel, dis);
}

// -------------------------------------------------------- void detection ---
// MarkVoidPoints with Advect's options (lsMarkVoidPoints.hpp:156-331): the
// face-connected regions of one sign (value >= 0 or < 0) of the frame box,
// by lock-free union-find: while the links are made, parents are read with
// device-scope atomic loads and written only by compare-and-swap (a root is
// only ever linked to a smaller root), so every XCD sees every link; a
// separate pass then points every entry at its root.  The top is the region
// of the last frame point with a value >= 0 in lexicographic order.
// q = (z ext1 + y) ext0 + x over the frame box.

__device__ inline uint32_t ufLoad(const uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ufRoot(const uint32_t *par, uint32_t a) {
  uint32_t p = ufLoad(&par[a]);
  while (p != a) {
    a = p;
    p = ufLoad(&par[a]);
  }
  return a;
}
__device__ inline void ufUnion(uint32_t *par, uint32_t a, uint32_t b) {
  a = ufRoot(par, a);
  b = ufRoot(par, b);
  while (a != b) {
    if (a < b) {
      const uint32_t t = a;
      a = b;
      b = t;
    }
    const uint32_t old = atomicCAS(&par[a], a, b);
    if (old == a) return;
    a = ufRoot(par, old);
    b = ufRoot(par, b);
  }
}

__device__ inline void voidCoords(const Frame &f, uint64_t q, int32_t r[3]) {
  r[0] = (int32_t)(q % (uint64_t)f.ext[0]);
  const uint64_t t = q / (uint64_t)f.ext[0];
  r[1] = (int32_t)(t % (uint64_t)f.ext[1]);
  r[2] = (int32_t)(t / (uint64_t)f.ext[1]);
}

__global__ void kVoidInit(BandView b, uint64_t N, uint8_t *sign, uint32_t *par, uint32_t *topQ) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < N; q += (uint64_t)gridDim.x * blockDim.x) {
    int32_t r[3];
    voidCoords(b.f, q, r);
    sign[q] = b.value(r[0] + b.f.o[0], r[1] + b.f.o[1], r[2] + b.f.o[2]) >= 0 ? 1 : 0;
    par[q] = (uint32_t)q;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *topQ = 0;
}
template <int D> __global__ void kVoidHook(Frame f, GridInfo g, uint64_t N, const uint8_t *sign, uint32_t *par) {
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < N; q += (uint64_t)gridDim.x * blockDim.x) {
    int32_t r[3];
    voidCoords(f, q, r);
    uint64_t stride = 1;
    for (int d = 0; d < D; ++d) {
      uint64_t n;
      bool has = true;
      if (r[d] + 1 < f.ext[d]) n = q + stride;
      else if (g.bc[d] == PERIODIC) n = q - (uint64_t)(f.ext[d] - 1) * stride;
      else has = false;
      if (has && sign[n] == sign[q]) ufUnion(par, (uint32_t)q, (uint32_t)n);
      stride *= (uint64_t)f.ext[d];
    }
  }
}
// (the links are final here: a thread reads either an original parent or a
// root another thread already stored, both on the path to the same root)
__global__ void kVoidFlatten(uint64_t N, const uint8_t *sign, uint32_t *par, uint32_t *topQ) {
  uint32_t best = 0;
  for (uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; q < N; q += (uint64_t)gridDim.x * blockDim.x) {
    par[q] = ufRoot(par, (uint32_t)q);
    if (sign[q]) best = (uint32_t)q + 1u;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t y = __shfl_down(best, o, 64);
    best = (y > best) ? y : best;
  }
  if ((threadIdx.x & 63) == 0 && best) atomicMax(topQ, best);
}
// Marker of the k-th defined entry (point id k).
template <int D>
__global__ void kVoidMarks(BandView b, const uint32_t *dlist, const uint32_t *ndef, const uint32_t *par,
                           const uint32_t *topQ, double *marks) {
  const uint32_t n = *ndef;
  const uint32_t tq = *topQ;
  const uint32_t top = tq ? par[tq - 1] : 0xFFFFFFFFu;
  auto label = [&](const int32_t c[3]) -> uint32_t {
    int64_t r[3], q = 0;
    for (int d = 0; d < 3; ++d) {
      r[d] = (int64_t)c[d] - b.f.o[d];
      if (r[d] < 0 || r[d] >= b.f.ext[d]) return 0xFFFFFFFEu;
    }
    q = (r[2] * b.f.ext[1] + r[1]) * b.f.ext[0] + r[0];
    return par[q];
  };
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    const uint32_t i = dlist[k];
    int32_t c[3];
    candCoords(b, i, b.pos[i], c);
    const double v = b.val[i];
    bool isVoid;
    if (v >= 0) {
      isVoid = label(c) != top;
    } else {
      isVoid = true;
      for (int s = 0; s < 2 * D && isVoid; ++s) {
        int32_t p[3] = {c[0], c[1], c[2]};
        const int d = s % D;
        p[d] = b.g.map(d, (int64_t)c[d] + (s < D ? 1 : -1));
        const double w = b.value(p[0], p[1], p[2]);
        if (signbit(w) == signbit(v)) continue;
        if (label(p) == top) isVoid = false;
      }
    }
    marks[k] = isVoid ? 1.0 : 0.0;
  }
}

void launchVoidMarks(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, uint8_t *sign,
                     uint32_t *par, uint32_t *topQ, const uint32_t *dlist, const uint32_t *ndef, double *marks) {
  const uint64_t N = (uint64_t)f.ext[0] * f.ext[1] * f.ext[2];
  BandView v = viewOf(b, f, g);
  const int gb = gridFor((size_t)N, 8192);
  kVoidInit<<<gb, TPB, 0, s>>>(v, N, sign, par, topQ);
  if (D == 2) kVoidHook<2><<<gb, TPB, 0, s>>>(f, g, N, sign, par);
  else kVoidHook<3><<<gb, TPB, 0, s>>>(f, g, N, sign, par);
  kVoidFlatten<<<gb, TPB, 0, s>>>(N, sign, par, topQ);
  if (D == 2) kVoidMarks<2><<<gridFor(b.cap), TPB, 0, s>>>(v, dlist, ndef, par, topQ, marks);
  else kVoidMarks<3><<<gridFor(b.cap), TPB, 0, s>>>(v, dlist, ndef, par, topQ, marks);
}

// ------------------------------------------------------ surface & Chamfer ---
// Surface points of a band: the zero crossing on every grid edge between a
// defined point a and its defined +e_d neighbour b of the other sign
// ((va > 0) != (vb > 0)), at a + t e_d with t = va / (va - vb) -- the node
// positions of a marching-cubes surface (lsToSurfaceMesh.hpp) before its node
// merging.  Items j = entry * D + d in entry order, so the crossings of an
// entry are contiguous: start[i] is the first of entry i (start[n] = total),
// and the band's own bitmap serves as the spatial index of the points.

template <int D> __device__ inline bool surfCrossing(const BandView &b, uint32_t i, int d, double *pt) {
  const double va = b.val[i];
  if (isUndef(va)) return false;
  int32_t c[3];
  candCoords(b, i, b.pos[i], c);
  if (c[d] >= b.g.hi[d]) return false;
  int32_t q[3] = {c[0], c[1], c[2]};
  q[d] += 1;
  const BandView::Loc l = b.locate(q[0], q[1], q[2]);
  if (!l.inside || !((b.bits[l.w] >> l.bt) & 1ull)) return false;  // no entry at the neighbour
  const double vb = b.val[b.entryOf(l, b.bits[l.w], b.pre[l.w])];
  if (isUndef(vb) || (va > 0) == (vb > 0)) return false;
  if (pt) {
    const double t = va / (va - vb);
    for (int k = 0; k < D; ++k) pt[k] = ((double)c[k] + (k == d ? t : 0.0)) * b.g.dx;
  }
  return true;
}

template <int D> __global__ __launch_bounds__(256) void kSurfCount(BandView b, const uint32_t *nptr, uint32_t *partials) {
  const uint32_t n = *nptr * D;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t acc = 0;
  for (uint32_t j = lo + threadIdx.x; j < hi; j += TPB) acc += surfCrossing<D>(b, j / D, (int)(j % D), nullptr) ? 1u : 0u;
  const uint32_t t = blockSum(acc);
  if (threadIdx.x == 0) partials[logicalBlock()] = t;
}
template <int D>
__global__ __launch_bounds__(256) void kSurfList(BandView b, const uint32_t *nptr, const uint32_t *partials,
                                                 double *pts, uint32_t *start, uint32_t *total) {
  const uint32_t ne = *nptr, n = ne * D;
  uint32_t lo, hi;
  blockChunk(n, lo, hi, 1);
  uint32_t off = blockOffset(partials);
  for (uint32_t t = lo; t < hi; t += TPB) {
    const uint32_t j = t + threadIdx.x;
    double pt[3];
    const bool hit = (j < hi) && surfCrossing<D>(b, j / D, (int)(j % D), pt);
    uint32_t tot;
    uint64_t wm;
    const uint32_t q = off + blockBallotScan(hit, tot, wm);
    if (j < hi && j % D == 0) start[j / D] = q;
    if (hit)
      for (int k = 0; k < D; ++k) pts[(size_t)q * D + k] = pt[k];
    off += tot;
  }
  if (logicalBlock() == gridDim.x - 1 && threadIdx.x == 0) {
    *total = off;
    start[ne] = off;
  }
}
void launchSurfacePoints(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, Workspace &ws,
                         double *pts, uint32_t *start, uint32_t *total) {
  BandView v = viewOf(b, f, g);
  if (D == 2) {
    kSurfCount<2><<<g_scanBlocks, TPB, 0, s>>>(v, b.n, ws.partials);
    kSurfList<2><<<g_scanBlocks, TPB, 0, s>>>(v, b.n, ws.partials, pts, start, total);
  } else {
    kSurfCount<3><<<g_scanBlocks, TPB, 0, s>>>(v, b.n, ws.partials);
    kSurfList<3><<<g_scanBlocks, TPB, 0, s>>>(v, b.n, ws.partials, pts, start, total);
  }
}

// Distance from every point of P to the nearest surface point of band Q:
// grid points of Q are visited in Chebyshev shells around the query's cell;
// after shell r every unvisited grid point's crossings lie at least r dx away.
template <int D>
__global__ void kNearest(const double *P, const uint32_t *nP, BandView q, const uint32_t *qStart, const double *qPts,
                         int maxShell, double *dist) {
  const uint32_t n = *nP;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double p[3] = {0, 0, 0};
    int32_t c[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
      p[d] = P[(size_t)k * D + d];
      c[d] = (int32_t)floor(p[d] / q.g.dx);
    }
    double best2 = __builtin_inf();
    // the surface points of the band entry at grid point c + o
    auto visit = [&](int ox, int oy, int oz) {
      const int32_t g[3] = {c[0] + ox, c[1] + oy, c[2] + oz};
      bool in = true;
      for (int d = 0; d < D; ++d) in = in && g[d] >= q.g.lo[d] && g[d] <= q.g.hi[d];
      if (!in) return;
      const BandView::Loc l = q.locate(g[0], g[1], g[2]);
      if (!l.inside || !((q.bits[l.w] >> l.bt) & 1ull)) return;
      const uint32_t e = q.entryOf(l, q.bits[l.w], q.pre[l.w]);
      for (uint32_t s = qStart[e]; s < qStart[e + 1]; ++s) {
        double d2 = 0;
        for (int d = 0; d < D; ++d) {
          const double t = p[d] - qPts[(size_t)s * D + d];
          d2 += t * t;
        }
        if (d2 < best2) best2 = d2;
      }
    };
    for (int r = 0; r <= maxShell; ++r) {
      // the grid points at Chebyshev distance r only: the faces of the
      // (2r+1)^D cube, each point once ((2r+1)^D - (2r-1)^D of them)
      if (r == 0) {
        visit(0, 0, 0);
      } else if (D == 2) {
        for (int ox = -r; ox <= r; ++ox) {
          visit(ox, -r, 0);
          visit(ox, r, 0);
        }
        for (int oy = -r + 1; oy <= r - 1; ++oy) {
          visit(-r, oy, 0);
          visit(r, oy, 0);
        }
      } else {
        for (int oy = -r; oy <= r; ++oy)
          for (int ox = -r; ox <= r; ++ox) {
            visit(ox, oy, -r);
            visit(ox, oy, r);
          }
        for (int oz = -r + 1; oz <= r - 1; ++oz) {
          for (int ox = -r; ox <= r; ++ox) {
            visit(ox, -r, oz);
            visit(ox, r, oz);
          }
          for (int oy = -r + 1; oy <= r - 1; ++oy) {
            visit(-r, oy, oz);
            visit(r, oy, oz);
          }
        }
      }
      // (one shell of margin: p / dx may round up to the next cell)
      const double rr = (double)(r - 1) * q.g.dx;
      if (r >= 1 && best2 <= rr * rr) break;
    }
    dist[k] = sqrt(best2);
  }
}
void launchNearest(hipStream_t s, int D, const double *P, const uint32_t *nP, size_t cap, const BandBufs &q,
                   const Frame &f, const GridInfo &g, const uint32_t *qStart, const double *qPts, int maxShell,
                   double *dist) {
  BandView v = viewOf(q, f, g);
  if (D == 2) kNearest<2><<<gridFor(cap), TPB, 0, s>>>(P, nP, v, qStart, qPts, maxShell, dist);
  else kNearest<3><<<gridFor(cap), TPB, 0, s>>>(P, nP, v, qStart, qPts, maxShell, dist);
}

// CompareVolume (lsCompareVolume.hpp:257-289): per cell of the box (lower
// corner lo + (x, y, z), extents ext), the corner averages of both level
// sets, corners summed in corner order and divided by 2^D; counts the cells
// inside (average <= 0) in exactly one of them.
template <int D>
__global__ void kCompareVolume(BandView A, BandView B, int64_t lo0, int64_t lo1, int64_t lo2, uint64_t e0,
                               uint64_t e1, uint64_t ncell, unsigned long long *count) {
  unsigned long long c = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < ncell;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const int64_t cell[3] = {lo0 + (int64_t)(k % e0), lo1 + (int64_t)((k / e0) % e1), lo2 + (int64_t)(k / (e0 * e1))};
    double ta = 0., tb = 0.;
    for (int i = 0; i < (1 << D); ++i) {
      int32_t p[3] = {0, 0, 0};
      for (int d = 0; d < D; ++d) p[d] = A.g.map(d, cell[d] + ((i >> d) & 1));
      ta += A.value(p[0], p[1], p[2]);
      tb += B.value(p[0], p[1], p[2]);
    }
    ta /= (1 << D);
    tb /= (1 << D);
    c += ((ta <= 0.) != (tb <= 0.)) ? 1u : 0u;
  }
  const unsigned long long t = blockSumU64(c);
  if (threadIdx.x == 0 && t) atomicAdd(count, t);
}
void launchCompareVolume(hipStream_t s, int D, const BandBufs &a, const Frame &fa, const BandBufs &b, const Frame &fb,
                         const GridInfo &g, const int64_t *lo, const int64_t *hi, unsigned long long *count) {
  const uint64_t e0 = (uint64_t)(hi[0] - lo[0] + 1), e1 = (uint64_t)(hi[1] - lo[1] + 1);
  const uint64_t e2 = (D == 3) ? (uint64_t)(hi[2] - lo[2] + 1) : 1u;
  const uint64_t n = e0 * e1 * e2;
  const BandView va = viewOf(a, fa, g), vb = viewOf(b, fb, g);
  if (D == 2) kCompareVolume<2><<<gridFor(n), TPB, 0, s>>>(va, vb, lo[0], lo[1], 0, e0, e1, n, count);
  else kCompareVolume<3><<<gridFor(n), TPB, 0, s>>>(va, vb, lo[0], lo[1], lo[2], e0, e1, n, count);
}

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
