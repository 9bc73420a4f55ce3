// Host-visible launch interface of the HIP kernels (lsg_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lsg_items.hpp"

namespace lsg {

// default grid of the band-operation scans: 1536 blocks measured best at
// the trench benchmark (1.479 ms per step against 1.506 at 1024, 1.683 at
// 2048, 2.70 at 256)
constexpr int SCAN_BLOCKS = 1536;
constexpr int SCAN_BLOCKS_MAX = 4096;
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 1;
// bitmap words per tile of a single-pass band operation
constexpr int BAND_TILE_WORDS = 32;

// Device scalars shared by the kernels of one advection handle.
struct DevScalars {
  unsigned long long maxAbsBits;  // max |H - D| over active points (bits of a non-negative double)
  unsigned long long alphaBits[3];  // LLF/LLLF per-direction alpha maxima; LF global alphas
  unsigned long long activeCount;   // active points of the last rate pass
  unsigned long long totalActive;   // accumulated active points (point updates)
  uint32_t overflow;
  uint32_t reframe;  // batch: an entry came close to its frame edge
  double dt[4];       // stage time steps (dt1, dt2, dt3) and a scratch slot
  double stepTime;    // time advanced by the last step
  double maxTimeStep; // bound passed to the first stage
  double rateBytes;   // accumulated algorithmic bytes of the rate kernel
  double starBytes;   // accumulated algorithmic bytes of the star-operation evaluation (single band)
  double star2Bytes;  // the same for the evaluation of two Expand cycles at once
  unsigned long long minStepBits;  // several level sets: min over points of the step limit (bits of a double >= 0)
  uint32_t movedDown;              // combine: the weighted value exceeded the start value somewhere
  unsigned long long stepActive0;  // totalActive at the start of the step
  // stencil local Lax-Friedrichs: ((j + 1) << 4) | material of the last
  // active point j (active-list order) whose scheme call set its alphas
  unsigned long long sllfKey;
  // single-pass band operations: the launch counter tagging tile status
  // words, and a flag set when a look-back wait ran out (never expected)
  uint32_t epoch;
  uint32_t lookbackFail;
};

// Mutable device buffers of a band.
struct BandBufs {
  uint64_t *bits;
  uint32_t *pre;
  double *val;
  uint32_t *pos;
  uint32_t *n;  // device entry count
  size_t cap;
  uint16_t *seg;  // batch bands only: level set of each entry
};

struct Workspace {
  uint64_t *cbits;     // nwords + 1
  uint64_t *cbits2;    // nwords + 1 (two Expand cycles: the first star dilation)
  uint32_t *cpre;      // nwords + 1
  uint32_t *cpos;      // ccap
  double *out;         // ccap
  uint32_t *ks;        // ccap + 1
  uint64_t *keepBits;  // ccap / 64 + 2 (keep flag per candidate)
  uint32_t *partials;  // SCAN_BLOCKS + 1
  uint32_t *ncand;     // device candidate count
  size_t ccap;
  uint16_t *cseg;      // ccap (batch: level set of each candidate)
  uint64_t *tileStat;  // nwords / BAND_TILE_WORDS + 1 (single-pass band operations: tile status)
};

// A batch launch: level-set table and per-word level set (device), words.
struct SegCtx {
  const SegInfo *segs;
  uint32_t nseg;
  uint32_t totalBase;
  const uint16_t *wseg;
  uint32_t nw;
};

BandView viewOf(const BandBufs &b, const Frame &f, const GridInfo &g);

// The single-pass band-operation kernels: whether they are used, and their
// switch-off after a look-back wait ran out (DevScalars::lookbackFail).
bool singlePassBandOps();
void disableSinglePassBandOps();
void setSinglePassBandOps(bool on);
void setScanBlocks(int blocks);
int defaultScanBlocks();
int scanBlocks();
void launchStarOpSeg(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const SegCtx &c, Workspace &ws,
                     const OpParams &op, DevScalars *sc);
void launchSelfOpSeg(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const SegCtx &c, Workspace &ws,
                     const OpParams &op, DevScalars *sc);
void launchRatesSeg(hipStream_t s, int D, const BandBufs &b, const SegCtx &c, const RateParams &rp, double *rates,
                    DevScalars *sc, unsigned long long *segMax, const double *maxDt, double *dtOut, Workspace &ws);
void launchStepBeginSeg(hipStream_t s, DevScalars *sc, double *maxDt, uint32_t nseg);
void launchStepTimeSeg(hipStream_t s, const DevScalars *sc, const double *dt, uint32_t nseg, int temporal,
                       double *stepTime, double *advTime);
void launchFrameCheck(hipStream_t s, const BandBufs &b, const Frame &f, const GridInfo &g, int D, int slack,
                      DevScalars *sc);
void launchFrameCheckSeg(hipStream_t s, const BandBufs &b, const SegCtx &c, int D, int slack, DevScalars *sc);
void launchCopyBandWords(hipStream_t s, const BandBufs &src, BandBufs &dst, int64_t nw);
void launchBooleanSeg(hipStream_t s, int D, const BandBufs &A, const BandBufs &B, BandBufs &out, const SegCtx &c,
                      Workspace &ws, int kind, double wTarget, double wSource, DevScalars *sc);

// Level sets below the advected one (several level sets), in the top's frame.
constexpr int MAX_LEVEL_SETS = 8;
struct LowerViews {
  int n;
  const uint64_t *bits[MAX_LEVEL_SETS - 1];
  const uint32_t *pre[MAX_LEVEL_SETS - 1];
  const double *val[MAX_LEVEL_SETS - 1];
};
// Rate pass with several level sets: per active point of the top a list of
// (velocity, stop value) pairs, (lw.n + 1) pairs per entry of `lists`; dt
// from the per-point step limits.  The learned scheme is not supported here.
void launchRatesMM(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const LowerViews &lw,
                   const RateParams &rp, double *lists, DevScalars *sc, const double *maxTimeStep, double *dtOut,
                   Workspace &ws);
// Advect::findGlobalAlphas with the material of each active point (the
// first level set, bottom up, whose value there is <= value + 1e-4) and its
// built-in functor (lsAdvect.hpp:90-195).
void launchGlobalAlphasMM(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g,
                          const LowerViews &lw, const RateParams &rp, DevScalars *sc);
// Records of host velocity calls per active point for a scheme: 1 (the
// scheme's call), 2 for LF (and findGlobalAlphas' call), 1 + 3^D for LLF
// (and the alpha box) and for the analytical LLF (and its box of
// getDissipationAlpha calls).
int hostVelocityRecords(int spatial, int D);
// Inputs of the host velocity calls: 8 doubles per record (coordinate,
// normal, point id, material or -1 for every material), records of active
// point j at [j * hostVelocityRecords, ...).
void launchVelInputs(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const RateParams &rp,
                     const LowerViews &lw, Workspace &ws, uint32_t *dlist, uint32_t *ndef, double *out,
                     DevScalars *sc);
// Second round of the stencil local Lax-Friedrichs host calls: records at
// the perturbed box normals (2 D per box point, 8 doubles each) from the
// first round's records `in` and results `vel`, for L materials.
void launchSllfInputs(hipStream_t s, int D, const uint32_t *nact, int L, const double *in, const double *vel,
                      double baseEps, double *out);
void launchSetFlag(hipStream_t s, uint32_t *flag, uint32_t v);

// Band operations (all asynchronous on `s`; sizes live in device memory).
// Star operations: e0/e1 (optional) are recorded around the evaluation
// kernel (kEvalKeep<EvalStar>), external when the stream is being captured.
void launchStarOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc, hipEvent_t e0 = nullptr,
                  hipEvent_t e1 = nullptr, bool externalEvents = false);
// Two Expand cycles as one band operation (multi-kernel pipeline, one level set).
void launchStarOp2(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                   Workspace &ws, const OpParams &op, DevScalars *sc, hipEvent_t e0 = nullptr,
                   hipEvent_t e1 = nullptr, bool externalEvents = false);
// updateLevelSet and rebuildLS as one star operation (multi-kernel
// pipeline, one level set; op.rates, op.dt, op.a = cutoff).
void launchUpdRebuildOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                        Workspace &ws, const OpParams &op, DevScalars *sc);
void launchSelfOp(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f, const GridInfo &g,
                  Workspace &ws, const OpParams &op, DevScalars *sc);
// movedDown (optional): set to 1 where a combine's weighted value exceeds
// the second band's value at a point defined in both (lsAdvect.hpp:218-224).
void launchBoolean(hipStream_t s, int D, const BandBufs &A, const BandBufs &B, BandBufs &out, const Frame &f,
                   const GridInfo &g, Workspace &ws, int kind, double wTarget, double wSource, DevScalars *sc,
                   uint32_t *movedDown = nullptr);
// Star operation with caller-provided scratch: the counters in *sc and the
// tile status words in stat (at least nwords / BAND_TILE_WORDS + 2).
void launchStarOpScratch(hipStream_t s, int D, const BandBufs &in, BandBufs &out, const Frame &f,
                         const GridInfo &g, const OpParams &op, uint64_t *stat, DevScalars *sc);
// Clears the overflow, look-back and re-frame flags of *sc.
void launchClearFlags(hipStream_t s, DevScalars *sc);
// Rate pass: Hamiltonian + dissipation per active entry, CFL reduction.
void launchRates(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, const RateParams &rp,
                 double *rates, DevScalars *sc, const double *maxTimeStep, double *dtOut, Workspace &ws,
                 hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, bool externalEvents = false);
// LF global alphas (Advect::findGlobalAlphas).
void launchGlobalAlphas(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g,
                        const RateParams &rp, DevScalars *sc);
// Small device-scalar kernels.
void launchStepTime(hipStream_t s, DevScalars *sc, int temporal);
// Step prologue: clear the overflow flag, mark the active-point counter and
// read maxTimeStep from host-mapped memory (so a captured step stays valid).
void launchStepBegin(hipStream_t s, DevScalars *sc, const double *maxDt);
void launchCopyBand(hipStream_t s, const BandBufs &src, BandBufs &dst, const Frame &f);
// Count defined entries (to *outCount) and the bounding box of in-frame
// entries (frame-relative, to bbox[6]: min x,y,z then max x,y,z).
// Entries before `first` are skipped.
void launchBandStats(hipStream_t s, const BandBufs &b, const Frame &f, uint32_t *outCount, int32_t *bbox,
                     uint32_t first = 0);
void launchFill32(hipStream_t s, uint32_t *p, uint32_t v, size_t n);
void launchNegate(hipStream_t s, const BandBufs &b);

// Point data: point ids of a band, the sources of an operation's result,
// one gather per array, the saved advection velocities.
void launchPointIds(hipStream_t s, const BandBufs &b, Workspace &ws, uint32_t *dlist, uint32_t *ndef,
                    uint32_t *rankOf);
void launchPdSources(hipStream_t s, int D, const BandBufs &out, const BandBufs &in, const BandBufs *inB, const Frame &f,
                     const GridInfo &g, const PdSrcParams &p, Workspace &ws, const uint32_t *rankIn,
                     const uint32_t *rankB, uint32_t *src, uint8_t *side, uint32_t *entry, uint32_t *nPush);
void launchPdGather(hipStream_t s, int comps, const double *A, const uint32_t *lenA, const double *B,
                    const uint32_t *lenB, const uint32_t *src, const uint8_t *side, const uint32_t *nPush, size_t cap,
                    double *out);
void launchCopyU32(hipStream_t s, const uint32_t *from, uint32_t *to);
// Surface points (zero crossings on grid edges) of a band: D doubles each,
// start[entry] = first point of the entry, *total = count.
void launchSurfacePoints(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, Workspace &ws,
                         double *pts, uint32_t *start, uint32_t *total);
// Nearest surface point of band q for each of the *nP points P.
void launchNearest(hipStream_t s, int D, const double *P, const uint32_t *nP, size_t cap, const BandBufs &q,
                   const Frame &f, const GridInfo &g, const uint32_t *qStart, const double *qPts, int maxShell,
                   double *dist);
// CompareVolume: cells of the box lo..hi (inclusive lower-corner indices)
// whose corner average is inside (<= 0) in exactly one of a, b, added to
// *count.
void launchCompareVolume(hipStream_t s, int D, const BandBufs &a, const Frame &fa, const BandBufs &b, const Frame &fb,
                         const GridInfo &g, const int64_t *lo, const int64_t *hi, unsigned long long *count);
// MarkVoidPoints: the marker (1 void, 0 not) of every defined point of b
// (dlist / ndef: the defined entries in order); sign / par hold one byte /
// one word per frame point.
void launchVoidMarks(hipStream_t s, int D, const BandBufs &b, const Frame &f, const GridInfo &g, uint8_t *sign,
                     uint32_t *par, uint32_t *topQ, const uint32_t *dlist, const uint32_t *ndef, double *marks);
void launchPdVelocities(hipStream_t s, const uint32_t *entry, const uint32_t *nPush, const double *val,
                        const double *rates, const double *diss, const double *dt, int nmat, size_t cap, double *vel,
                        double *dis);

}  // namespace lsg

namespace lsg {
void launchWeno(hipStream_t s, int order, uint32_t n, const double *x, double dx, int plus, double *out);
}
