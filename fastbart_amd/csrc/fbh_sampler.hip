// The batched BART sampler on gfx950: one persistent workgroup per chain runs whole sweeps on the device.
//
// Per tree update (src/bart_kernel.cpp:400-535): proposal, one data pass over the chain's rows (apply the previous
// tree's accepted change to the leaf ids and the running fit, then this tree's sufficient statistics), the
// Metropolis-Hastings decision and the leaf draws. After the last tree, a closing pass (apply, residual sum of squares,
// posterior-mean sum) and the sigma draw or the probit latent redraw (:536-558).
//
// Workgroup = 1 control wave + 4 data waves.
//  * Data wave w owns summation lanes 4w..4w+3 (row i belongs to lane i % 16, src/bart_kernel.cpp:354-356). It runs
//    phase A (apply + per-row w, w*r and a tag = leaf rank | birth-split side) for the rows of its own lanes and then
//    phase B, the ordered per-(lane, leaf) sums: one thread per (lane, task) walks its lane's rows in row order with a
//    fused multiply-add by 1.0 or 0.0 (adding +0.0 to a sum that started at +0.0 changes nothing, so these are exactly
//    the reference's lane sums). A wave only ever reads rows it wrote itself, so the pass needs no workgroup barrier.
//    In FX64 mode the same walk adds int64 fixed-point images instead (order free, exact).
//  * The control wave overlaps the data pass of tree j with: the 1 + 2*NODES generator outputs of tree j's MH test and
//    leaf normals, the libm-image gathers for the slots that can be leaves after the MH step, and the proposal of tree
//    j+1 (whose draws follow those in the stream). After the pass it adds the 16 lanes, runs MH and the leaf draws.
//  * The running fit of a chain stays in LDS across the whole launch when it fits (RES); leaf ids, cutpoint ranks, y and
//    w stream from HBM / L2 with all of a tile's loads issued before use.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

constexpr int kDataWaves = 4;
constexpr int kBlock = 64 * (1 + kDataWaves);
constexpr int kLanesPerWave = kLanes / kDataWaves;   // 4 summation lanes per data wave
constexpr int kTaskSlots = 64 / kLanesPerWave;       // 16 task threads per lane
constexpr int kMaxTasks = 2 * kMaxRank;              // (sum of w, sum of w*r) for every leaf and split half
constexpr int kGroups = (kMaxTasks + kTaskSlots - 1) / kTaskSlots;
constexpr int kTreeWords = sizeof(TreeDev) / 8;  // 90

// Tile geometry: R rows per phase-A thread per tile, 16 R rows per lane, 256 R rows per tile.
template <int R>
struct Geo {
  static constexpr int kTW4 = 16 * R;                         // rows of one lane in a tile
  static constexpr int kTile = kLanes * kTW4;                 // rows in a tile
  static constexpr int kStride = kTW4 + 1;                    // doubles per lane row block (bank offset)
  static constexpr int kTagStride = (kTW4 + 4) & ~3;          // bytes per lane tag block
  // the w*r blocks start 8 doubles (16 banks) after the w blocks end: a phase-B wave reads w and w*r of the same row of
  // the same lane in one instruction (even and odd task slots), which then fall in different banks
  static constexpr int kV1Pad = 8;
  static constexpr int kScratchBytes = (2 * kLanes * kStride + kV1Pad) * 8 + kLanes * kTagStride;
};

struct KArgs {
  const ChainDesc* desc;
  ChainState* state;
  TreeDev* trees;
  const uint8_t* q;
  double* Y;           // probit latents
  const float* yf;     // wbart outcome
  const float* W;
  double* AF;
  double* TM;
  const uint8_t* ybin;
  uint8_t* leaf;
  float* sigma_out;
  uint64_t* trees_out;
  const double* PG;
  LibmTables tb;
  TraceRec* trace;
  int trace_cap;  // records per chain
  int m, nskip, ndpost, probit, keep_trees, stats_mode, node_cap;
  int it0, it1;
  int af_off;     // byte offset of the resident running fit in dynamic LDS
  unsigned long long* xbuf;  // row split: [group][parity][member][kXw] pass totals
  uint32_t* xctr;            // row split: [group] arrivals
  const int* order;          // [C] device chain run by workgroup w of the sweep kernel
  int pr_ext;                // probit: the latents are redrawn by the k_pr_* kernels after the launch (one sweep)
};

struct Rng {
  uint32_t s0, s1, s2, s3;
};

// the sweep kernel's dynamic LDS: tile scratch (also the probit ring), then the resident running fit
extern __shared__ __align__(16) unsigned char fbh_dyn[];

__device__ __forceinline__ uint32_t ufl(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
// a wave-uniform 64-bit value (pointer or double) read from the first lane
template <class T>
__device__ __forceinline__ T ufl_t(T v) {
  static_assert(sizeof(T) == 8, "64-bit values");
  uint64_t u;
  __builtin_memcpy(&u, &v, 8);
  u = ((uint64_t)ufl((uint32_t)(u >> 32)) << 32) | ufl((uint32_t)u);
  __builtin_memcpy(&v, &u, 8);
  return v;
}
// loads and stores through pointers known to be in global memory (generic pointers passed to a function that is not
// inlined would otherwise become flat accesses, which also count against the LDS wait counter)
template <class T>
__device__ __forceinline__ T gld(const T* p) {
  typedef const __attribute__((address_space(1))) T* gptr;
  return *(gptr)p;
}
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
  typedef __attribute__((address_space(1))) T* gptr;
  *(gptr)p = v;
}
// element i of a global array whose base is wave-uniform: a 32-bit byte offset from the scalar base, so that no 64-bit
// per-row address is formed (nor kept live across loops)
template <class T>
__device__ __forceinline__ T ldo(const T* base, uint32_t i) {
  typedef const __attribute__((address_space(1))) T* gptr;
  return *(gptr)(reinterpret_cast<const char*>(base) + (size_t)(i * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void sto(T* base, uint32_t i, T v) {
  typedef __attribute__((address_space(1))) T* gptr;
  *(gptr)(reinterpret_cast<char*>(base) + (size_t)(i * (uint32_t)sizeof(T))) = v;
}
// a per-thread value that the compiler may not hoist out of a loop: the row indices derived from it are recomputed
// where they are used (those of the first tile are the same in every pass; kept live across the sweep they would be
// spilled, and a reload from scratch waits for every load issued before it)
__device__ __forceinline__ uint32_t opaque(uint32_t i) {
  asm volatile("" : "+v"(i));
  return i;
}

// Index of row i in the LDS-resident running fit: one pad slot after every 16 rows. A phase-A wave touches rows
// 16 s + k for 16 consecutive s and its 4 lanes k; unpadded, those doubles fall in 4 bank pairs (16-way conflicts),
// padded in all 16.
__device__ __forceinline__ uint32_t afl(uint32_t i) { return i + (i >> 4); }
inline size_t afl_doubles(size_t n) { return n + (n + 15) / 16; }

// xoshiro128** (Blackman & Vigna), state order of src/bart_kernel.cpp:54-58
__device__ __forceinline__ uint32_t rng_next(Rng& g) {
  const uint32_t x = g.s1 * 5u;
  const uint32_t res = ((x << 7) | (x >> 25)) * 9u;
  const uint32_t t = g.s1 << 9;
  g.s2 ^= g.s0;
  g.s3 ^= g.s1;
  g.s1 ^= g.s2;
  g.s0 ^= g.s3;
  g.s2 ^= t;
  g.s3 = (g.s3 << 11) | (g.s3 >> 21);
  return res;
}
// the two halves of rng_next: the output from s1 before the step, and the state step alone. k11 / k21 are 11 and 21
// passed at run time so that the rotation is not matched as one (it then stays on the scalar unit).
__device__ __forceinline__ uint32_t starstar(uint32_t s1) {
  const uint32_t x = s1 * 5u;
  return ((x << 7) | (x >> 25)) * 9u;
}
__device__ __forceinline__ void rng_step(Rng& g, uint32_t k11, uint32_t k21) {
  const uint32_t t = g.s1 << 9;
  g.s2 ^= g.s0;
  g.s3 ^= g.s1;
  g.s1 ^= g.s2;
  g.s0 ^= g.s3;
  g.s2 ^= t;
  g.s3 = (g.s3 << k11) | (g.s3 >> k21);
}
// (k + 1/2) 2^-24 from the top 24 bits (:59)
__device__ __forceinline__ float unif_f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
// sqrtf(-2 logf(u1)) cosf(2 pi u2) in float with the host libm images (:60). __builtin_sqrtf / __builtin_sqrt lower to
// correctly rounded sequences (v_sqrt + FMA residual correction); HIP's __fsqrt_rn is the native approximation unless
// OCML_BASIC_ROUNDED_OPERATIONS is defined, so it is not used anywhere.
__device__ __forceinline__ float normal_f(uint32_t x1, uint32_t x2, const LibmTables& tb) {
  return __builtin_sqrtf(-2.f * gld(tb.logf_t + (x1 >> 8))) * gld(tb.cosf_t + (x2 >> 8));
}
__device__ __forceinline__ double unif_d(Rng& g) { return (double)unif_f(rng_next(g)); }
__device__ __forceinline__ double rnorm_d(Rng& g, const LibmTables& tb) {
  const uint32_t a = rng_next(g);
  const uint32_t b = rng_next(g);
  return (double)normal_f(a, b, tb);
}
// log((double)u) of the next uniform, from the host-libm image
__device__ __forceinline__ double log_unif(Rng& g, const LibmTables& tb) { return tb.logd_t[rng_next(g) >> 8]; }

// Marsaglia-Tsang gamma, shape >= 1 (:73-83)
__device__ double rgamma_dev(Rng& g, double a, const LibmTables& tb) {
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / __builtin_sqrt(9.0 * d);
  for (int tries = 0; tries < 32; tries++) {
    const double x = rnorm_d(g, tb);
    double v = 1.0 + c * x;
    if (v <= 0) continue;
    v = v * v * v;
    const double lu = log_unif(g, tb);
    const double rhs = 0.5 * x * x + d - d * v + d * log(v);
    if (lu < rhs) return d * v;
  }
  return d;
}

// truncated normal above tau (:282-292)
__device__ double rtnorm_dev(Rng& g, double mean, double tau, const LibmTables& tb) {
  double z;
  tau = tau - mean;
  if (tau <= 0.) {
    do {
      z = rnorm_d(g, tb);
    } while (z < tau);
  } else {
    const double lam = 0.5 * (tau + __builtin_sqrt(tau * tau + 4.0));
    for (;;) {
      const double e = -log_unif(g, tb);
      z = e / lam + tau;
      const double u = unif_d(g);
      if (!(u > exp(-0.5 * (z - lam) * (z - lam)))) break;
    }
  }
  return z + mean;
}

// ------------------------------------------------------------------ probit latent redraw in windows
// Every attempt of rtnorm (:282-292) consumes exactly two generator outputs -- one attempt *slot*: (u1, u2) for one
// normal when tau' <= 0, or (u for -log u, u for the test) when tau' > 0. Row i's latent is rtnorm(f, -off) if y = 1,
// -rtnorm(-f, off) if y = 0 (:543-547). Rows go in windows of kPR starting at slot S: for every (row l, offset t) of
// the window it is decided whether an attempt of row l in slot S + l + t would accept; row l's attempts are the slots
// S + l + r_l, S + l + r_l + 1, ... where r_l is the number of rejections of the rows before it in the window, so a
// scan of the decision masks finds each row's accepting slot. Decisions are taken in float (double for tau' <= 0's
// normal) with a guard band and fall back to the exact double / host-libm-image evaluation whenever the two sides of the
// comparison are close enough that rounding could matter, so every decision is the reference's; the accepted draws are
// then evaluated exactly from the images.
constexpr int kPR = 64;             // rows per window
constexpr int kBand = 16;           // attempt offsets decided per row: [b_l, b_l + kBand)
constexpr int kRhoMax = 24;         // cap of the predicted rejections per row, in 1/16
constexpr int kTMax = (kRhoMax * (kPR - 1)) / 16 + kBand;  // largest offset a band reaches
constexpr int kNeed = kPR + kTMax;  // slots past a window's first slot its decisions read
constexpr int kRing = 512;          // generated slots kept in LDS (generated 64 at a time)
constexpr int kCkpt = 8;            // a generator-state checkpoint every kCkpt slots
constexpr int kAhead = 256;         // slots generated past a window's first slot while its decisions run
static_assert(kAhead >= kNeed && kRing >= kAhead + kNeed, "ring too small for the band");
// first attempt offset decided for row l of a window: the predicted rejections before it (rho16 / 16 per row) less a
// margin, so that the true offset falls inside [b_l, b_l + kBand) unless the window's rows deviate by more than kBand/2
__device__ __forceinline__ int band_start(int l, int rho16) { return max(0, (rho16 * l) / 16 - kBand / 2); }

struct ProbitRing {
  uint32_t raw[kRing][2];         // s1 before each of slot k's two outputs (k mod kRing); the output is starstar(s1)
  uint32_t ck[kRing / kCkpt][4];  // generator state before slot kCkpt * q
  uint32_t dmask[kPR];            // row l: bit b set if the attempt in slot S + l + band_start(l) + b accepts
  uint8_t acc[kPR];               // accepting offset of each resolved row
  int p_done, p_used;             // rows resolved by the window, slots they used
  uint32_t fin[4];                // generator state after the last row
};
static_assert(kBand == 16 && kPR == 64, "four rows of 16 offsets per wave-wide decision");

struct LatentRow {
  double mean, tau, lam;          // rtnorm's mean, tau - mean, and the exponential rate when tau - mean > 0
  float tauf, lamf;
  bool pos;                       // y = 1: z = rtnorm(f, -off); y = 0: z = -rtnorm(-f, off)
};

__device__ __forceinline__ LatentRow latent_row(double f, int yb, double off) {
  LatentRow r;
  r.pos = yb != 0;
  r.mean = r.pos ? f : -f;
  const double tau_in = r.pos ? -off : off;
  r.tau = tau_in - r.mean;
  r.lam = r.tau > 0. ? 0.5 * (r.tau + __builtin_sqrt(r.tau * r.tau + 4.0)) : 0.;
  r.tauf = (float)r.tau;
  r.lamf = (float)r.lam;
  return r;
}

// exact decision of one attempt (the reference's arithmetic; image gathers)
__device__ __forceinline__ bool latent_accepts_exact(const LatentRow& r, uint32_t x1, uint32_t x2,
                                                     const LibmTables& tb) {
  if (r.tau <= 0.) return (double)normal_f(x1, x2, tb) >= r.tau;  // while (z < tau) retry
  const double u = (double)unif_f(x2);
  const double e = -gld(tb.logd_t + (x1 >> 8));
  const double z = e / r.lam + r.tau;
  return !(u > exp(-0.5 * (z - r.lam) * (z - r.lam)));
}

// the same decision, from device math where it is unambiguous
__device__ __forceinline__ bool latent_accepts(const LatentRow& r, uint32_t x1, uint32_t x2, const LibmTables& tb) {
  if (r.tau <= 0.) {
    // logf / cosf / sqrtf within a few ulp of the host's: |z - z_host| < 1e-6 |z|
    const float za = sqrtf(-2.f * logf(unif_f(x1))) * cosf(6.28318531f * unif_f(x2));
    const double d = (double)za - r.tau;
    if (fabs(d) > 1e-5 * fabs((double)za) + 1e-30) return d >= 0.;
  } else if (r.tau < 64.) {
    // float evaluation of u > exp(-(z - lam)^2 / 2): relative error of the exponential below 1e-4 for tau < 64
    const float u = unif_f(x2);
    const float za = -logf(unif_f(x1)) / r.lamf + r.tauf;
    const float dz = za - r.lamf;
    const float ea = expf(-0.5f * dz * dz);
    if (fabsf(u - ea) > 1e-3f * u) return !(u > ea);
  }
  return latent_accepts_exact(r, x1, x2, tb);
}

// the latent of an accepted attempt, as rtnorm returns it and the caller signs it
__device__ __forceinline__ double latent_value(const LatentRow& r, uint32_t x1, uint32_t x2, const LibmTables& tb) {
  double z;
  if (r.tau <= 0.) {
    z = (double)normal_f(x1, x2, tb);
  } else {
    const double e = -gld(tb.logd_t + (x1 >> 8));
    z = e / r.lam + r.tau;
  }
  const double v = z + r.mean;
  return r.pos ? v : -v;
}

// ------------------------------------------------------------------ control state of one chain's workgroup
struct ThR {  // a slot's leaf value and, for a used leaf, its code (a row's task tag, leaf_code below)
  double th;
  uint32_t rank, pad_;
};

// A used leaf's code in the row tags: its rank among the tree's leaves, or, in the one-chain-per-CU build with REF16
// statistics (oh) for a tree of at most kOneHotLeaves leaves, one bit at the rank, which lets a summing thread test
// its task with a single bit-field extract (bits 6 and 7 of the tag are the birth proposal's halves). The dense build
// and FX64 statistics keep the ranks: measured, the one-bit loop is slower there.
constexpr int kOneHotLeaves = 6;
__device__ __forceinline__ uint32_t leaf_code(int rank, int nleaf, bool oh) {
  return (oh && nleaf <= kOneHotLeaves) ? (1u << rank) : (uint32_t)rank;
}
struct Ctl {
  TreeDev tree[2];               // tree j (being updated) and tree j+1 (proposal prepared during pass j)
  ChainDesc cd;
  double PG[kNodes];
  ThR thr[2][kNodes];            // leaf value and leaf rank of every slot of the tree whose statistics the pass computes
  double tho[kNodes], thn[kNodes];  // previous tree's values before / after its update (applied by the next pass)
  double S0[kNodes], S1[kNodes];
  // per-lane sums (double bits, or int64 in FX64 mode); a pad column so that the combine's lanes (one task each,
  // reading its 16 lanes in order) fall in different LDS banks
  double part[kMaxTasks][kLanes + 1];
  uint64_t cand[2];              // slots that can be leaves after the MH step
  double prr[2], lpr[2];           // proposal ratio and its log (taken while the data waves run)
  int cmode[2], cnx[2], cv[2], cc[2], nleaf[2], ngood[2], nnogs[2];
  uint8_t slot_of[2][kNodes];
  uint32_t rnd[2 * kNodes + 4];  // generator outputs of the MH uniform and the leaf normals
  uint32_t rng_mark[4];          // generator state right after them (trace records)
  float z[kNodes];
  double lu;
  int pmode, pnx, pv, pc, pL, pR, pl, pr;
  int nL, nR;
  double sL0, sL1, sR0, sR1, rss;
  double inv_s2;
  uint32_t rot[2];               // 11 and 21 (see rng_step)
  uint32_t pst_[4];              // generator state handed to the probit latent redraw
};

#define BITM(mask, i) (((mask) >> (i)) & 1ull)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// index of the k-th set bit (0-based); 0 if there are not k+1 bits (:94-101)
__device__ __forceinline__ int select_kth(uint64_t m, int k) {
  for (int i = 0; i < k && m; i++) m &= m - 1;
  return m ? __builtin_ctzll(m) : 0;
}
// lowest unused slot in [1, cap) other than skip, or -1 (:102-109)
__device__ __forceinline__ int first_free(uint64_t used, int skip, int cap) {
  uint64_t f = ~used & ~1ull;
  if (cap < 64) f &= (1ull << cap) - 1ull;
  if (skip >= 0) f &= ~(1ull << skip);
  return f ? __builtin_ctzll(f) : -1;
}
// BART's tree::rg (:238-251), uniform walk
__device__ void node_range(const TreeDev& t, int x, int v, int nc, int* L, int* U) {
  int lo = 0, hi = nc - 1, node = x;
  for (int s = 0; s < kMaxDepth && node != 0; s++) {
    const int pa = t.par[node];
    if ((int)t.v[pa] == v) {
      if ((int)t.l[pa] == node) {
        if ((int)t.c[pa] - 1 < hi) hi = (int)t.c[pa] - 1;
      } else {
        if ((int)t.c[pa] + 1 > lo) lo = (int)t.c[pa] + 1;
      }
    }
    node = pa;
  }
  *L = lo;
  *U = hi;
}
// variables with a non-empty range at node x (:253-276): lane v tracks variable v along the shared ancestor walk
__device__ int count_goodvars_wave(const TreeDev& t, int x, const int* nc, int p, int lane) {
  int lo = 0, hi = (lane < FBH_PMAX ? nc[lane] : 0) - 1;
  int node = x;
  for (int s = 0; s < kMaxDepth && node != 0; s++) {
    const int pa = t.par[node], v = t.v[pa], c = t.c[pa];
    if (v == lane) {
      if ((int)t.l[pa] == node) {
        if (c - 1 < hi) hi = c - 1;
      } else {
        if (c + 1 > lo) lo = c + 1;
      }
    }
    node = pa;
  }
  return __popcll(__ballot(lane < p && hi >= lo));
}
__device__ __forceinline__ double lh(double b, double M, double t2) {
  const double k = b * t2 + 1;
  return -0.5 * log(k) + 0.5 * M * M * t2 / k;
}

__device__ __forceinline__ uint64_t pack_word(const TreeDev& t, int s) {
  const uint32_t thb = __float_as_uint(t.th[s]);
  const uint64_t vv = BITM(t.used, s) ? (uint64_t)t.v[s] : 255u;
  return vv | ((uint64_t)t.c[s] << 8) | ((uint64_t)t.l[s] << 16) | ((uint64_t)t.r[s] << 24) | ((uint64_t)thb << 32);
}

__device__ __forceinline__ void copy_tree(TreeDev* dst, const TreeDev* src, int lane) {
  const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
  uint64_t* d = reinterpret_cast<uint64_t*>(dst);
  for (int w = lane; w < (int)(sizeof(TreeDev) / 8); w += 64) d[w] = s[w];
}

// ------------------------------------------------------------------ control wave pieces
// Proposal for tree buffer b (:413-464). All 64 lanes run it (lane s looks at slot s); the generator state and every
// decision are wave-uniform.
__device__ void ctl_propose(Ctl& c, int b, Rng& g, int cap, int lane, bool oh) {
  const TreeDev& t = c.tree[b];
  const double pb = 0.5;
  const int p = c.cd.p;
  const uint64_t used = t.used, leafm = t.leaf;
  const int s = lane;
  const bool isleaf = BITM(leafm, s), isused = BITM(used, s);
  const bool good = isused && isleaf && t.ngv[s] > 0;
  const bool nog = isused && !isleaf && BITM(leafm, t.l[s] & 63) && BITM(leafm, t.r[s] & 63);
  const uint64_t goodm = __ballot(good), nogm = __ballot(nog);
  const uint64_t lm = used & leafm;
  c.thr[b][s].th = (double)t.th[s];
  if (BITM(lm, s)) {
    const int rk = __popcll(lm & ((1ull << s) - 1ull));
    c.thr[b][s].rank = leaf_code(rk, __popcll(lm), oh);
    c.slot_of[b][rk] = (uint8_t)s;
  } else {
    c.thr[b][s].rank = 0;
  }
  const int ngood = __popcll(goodm), nnogs = __popcll(nogm);
  const bool single = used == 1ull;
  const double PBx = ngood == 0 ? 0 : (single ? 1 : pb);
  int cmode = 0, cnx = 0, cv = 0, cc = 0;
  double pr = 0;
  uint64_t cand = lm;
  if (unif_d(g) < PBx) {
    const int nx = select_kth(goodm, (int)(unif_d(g) * ngood));
    int v = (int)(unif_d(g) * p);
    if (v >= p) v = p - 1;
    int bL, bU;
    node_range(t, nx, v, c.cd.nc[v], &bL, &bU);
    if (bU >= bL && (int)t.d[nx] + 1 < kMaxDepth) {
      int cc_ = bL + (int)(unif_d(g) * (bU - bL + 1));
      if (cc_ > bU) cc_ = bU;
      const int dnx = t.d[nx];
      const double PGnx = c.PG[dnx];
      double PGly, PGry;
      if (t.ngv[nx] > 1) {
        PGly = PGry = c.PG[dnx + 1];
      } else {
        PGly = (cc_ - 1 < bL) ? 0 : c.PG[dnx + 1];
        PGry = (bU < cc_ + 1) ? 0 : c.PG[dnx + 1];
      }
      const double PDy = (ngood > 1) ? 1 - pb : ((PGry == 0 && PGly == 0) ? 1 : 1 - pb);
      double Pnogy;
      if (nx == 0) {
        Pnogy = 1;
      } else {
        const int pa = t.par[nx];
        const bool panog = BITM(leafm, t.l[pa] & 63) && BITM(leafm, t.r[pa] & 63);
        Pnogy = panog ? 1.0 / nnogs : 1.0 / (nnogs + 1);
      }
      const double Pbotx = 1.0 / ngood;
      pr = (PGnx * (1 - PGly) * (1 - PGry) * PDy * Pnogy) / ((1 - PGnx) * Pbotx * PBx);
      cmode = 1;
      cnx = nx;
      cv = v;
      cc = cc_;
      const int L = first_free(used, -1, cap), R = L < 0 ? -1 : first_free(used, L, cap);
      if (L >= 0 && R >= 0) cand |= (1ull << L) | (1ull << R);
    }
  } else {
    const int nx = select_kth(nogm, (int)(unif_d(g) * nnogs));
    const int dny = t.d[nx], l = t.l[nx], r = t.r[nx];
    const double PGny = c.PG[dny];
    const double PGlx = t.ngv[l] > 0 ? c.PG[dny + 1] : 0, PGrx = t.ngv[r] > 0 ? c.PG[dny + 1] : 0;
    const double PBy = (nx == 0) ? 1 : pb;
    const int ng2 = ngood - (t.ngv[l] > 0) - (t.ngv[r] > 0) + 1;
    const double Pboty = 1.0 / ng2, PDx = 1 - PBx, Pnogx = 1.0 / nnogs;
    pr = ((1 - PGny) * PBy * Pboty) / (PGny * (1 - PGlx) * (1 - PGrx) * PDx * Pnogx);
    cmode = 2;
    cnx = nx;
    cand |= 1ull << nx;
  }
  const double lpr = cmode ? log(pr) : 0.0;
  if (lane == 0) {
    c.cmode[b] = cmode;
    c.lpr[b] = lpr;
    c.cnx[b] = cnx;
    c.cv[b] = cv;
    c.cc[b] = cc;
    c.prr[b] = pr;
    c.cand[b] = cand;
    c.nleaf[b] = __popcll(lm);
    c.ngood[b] = ngood;
    c.nnogs[b] = nnogs;
  }
}

// Add the 16 lanes of every task in lane order from zero (:202-212).
template <bool FX>
__device__ void ctl_combine(Ctl& c, int b, bool rss, int cmode, int lane) {
  const int nleaf = c.nleaf[b];
  const int ntask = rss ? 2 : 2 * (nleaf + (cmode == 1 ? 2 : 0));
  c.S0[lane] = 0;
  c.S1[lane] = 0;
  if (lane == 0) c.sL0 = c.sL1 = c.sR0 = c.sR1 = c.rss = 0;
  wave_sync();
  for (int task = lane; task < ntask; task += 64) {
    double x;
    if (FX) {
      long long s = 0;
      for (int k = 0; k < kLanes; k++) s += __double_as_longlong(c.part[task][k]);
      x = (double)s * (rss ? (1.0 / 16777216.0) : (1.0 / 4294967296.0));
    } else {
      x = 0;
      for (int k = 0; k < kLanes; k++) x += c.part[task][k];
    }
    const int q = task >> 1, kind = task & 1;
    if (rss) {
      if (kind) c.rss = x;
    } else if (q < nleaf) {
      const int s = c.slot_of[b][q];
      if (kind) c.S1[s] = x;
      else c.S0[s] = x;
    } else if (q == nleaf) {
      if (kind) c.sL1 = x;
      else c.sL0 = x;
    } else {
      if (kind) c.sR1 = x;
      else c.sR0 = x;
    }
  }
  wave_sync();
}

// ------------------------------------------------------------------ data pass (data waves)
struct RowPtrs {
  double* AF;        // global running fit
  double* AFl;       // LDS running fit (resident mode)
  double* TM;
  const double* Y;
  const float* yf;
  const float* W;
  const uint8_t* Q;
  uint8_t* LEAF;
  int n, probit;
  int ns;  // column stride of Q and LEAF
};

// Walk one lane's rows in row order, adding src[i] to the task's sum where the task matches the row's tag: bit f of the
// task's pattern, f being the tag's field (sh, wd) -- the leaf rank (bits 0-4) or the birth-split side (bits 6-7); with
// one-bit leaf codes (OH), bit sh of the tag itself. Bit-field extracts and a conversion make the 0.0 / 1.0 factor: no
// compare, so no lane mask and no wait states on the sequential adds. Rows go in blocks of 16 whose LDS reads are all
// issued before the adds.
template <bool OH>
__device__ __forceinline__ uint32_t tag_match(uint32_t word, uint32_t off, uint32_t wd, uint32_t pat) {
  if (OH) return __builtin_amdgcn_ubfe(word, off, 1u);
  return __builtin_amdgcn_ubfe(pat, __builtin_amdgcn_ubfe(word, off, wd), 1u);
}
template <bool FX, bool OH>
__device__ __forceinline__ double lane_sum(const uint8_t* __restrict__ tag, const double* __restrict__ src, int cnt,
                                           uint32_t sh, uint32_t wd, uint32_t pat, double acc) {
  int i = 0;
  if (FX) {
    long long s = __double_as_longlong(acc);
    const long long* v = reinterpret_cast<const long long*>(src);
    for (; i + 16 <= cnt; i += 16) {
      uint32_t t4[4];
      long long x[16];
#pragma unroll
      for (int u = 0; u < 4; u++) t4[u] = *reinterpret_cast<const uint32_t*>(tag + i + 4 * u);
#pragma unroll
      for (int u = 0; u < 16; u++) x[u] = v[i + u];
#pragma unroll
      for (int u = 0; u < 16; u++) s += tag_match<OH>(t4[u >> 2], 8 * (u & 3) + sh, wd, pat) ? x[u] : 0ll;
    }
    for (; i < cnt; i += 8) {
      // the last rows in blocks of 8: rows at or past cnt hold earlier values and are masked out
      const int rem = cnt - i;
      uint32_t t2[2];
      long long x[8];
#pragma unroll
      for (int u = 0; u < 2; u++) t2[u] = *reinterpret_cast<const uint32_t*>(tag + i + 4 * u);
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = v[i + u];
#pragma unroll
      for (int u = 0; u < 8; u++) s += (u < rem && tag_match<OH>(t2[u >> 2], 8 * (u & 3) + sh, wd, pat)) ? x[u] : 0ll;
    }
    return __longlong_as_double(s);
  }
  double s = acc;
  for (; i + 16 <= cnt; i += 16) {
    uint32_t t4[4];
    double x[16];
#pragma unroll
    for (int u = 0; u < 4; u++) t4[u] = *reinterpret_cast<const uint32_t*>(tag + i + 4 * u);
#pragma unroll
    for (int u = 0; u < 16; u++) x[u] = src[i + u];
#pragma unroll
    for (int u = 0; u < 16; u++) s = fma((double)tag_match<OH>(t4[u >> 2], 8 * (u & 3) + sh, wd, pat), x[u], s);
  }
  for (; i < cnt; i += 8) {
    // the last rows in blocks of 8: rows at or past cnt hold earlier values (any bit pattern) and add +0.0
    const int rem = cnt - i;
    uint32_t t2[2];
    double x[8];
#pragma unroll
    for (int u = 0; u < 2; u++) t2[u] = *reinterpret_cast<const uint32_t*>(tag + i + 4 * u);
#pragma unroll
    for (int u = 0; u < 8; u++) x[u] = src[i + u];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const bool in = u < rem;
      s = fma(in ? (double)tag_match<OH>(t2[u >> 2], 8 * (u & 3) + sh, wd, pat) : 0.0, in ? x[u] : 0.0, s);
    }
  }
  return s;
}

__device__ __forceinline__ long long fx_round(double x, double scale) { return (long long)rint(x * scale); }

typedef float f4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));
struct D4 {  // four consecutive doubles
  d2v a, b;
};
__device__ __forceinline__ double d4_at(const D4& v, int u) { return u < 2 ? v.a[u] : v.b[u - 2]; }

// One phase-A thread's loads of a tile. Thread t of data wave dw takes the row slots t + 64 s of the tile (S of them)
// and, in each, the rows of the wave's four summation lanes -- four consecutive rows, so one dword of every byte column,
// one 16-byte vector of every float column and two of every double column (columns are 16-byte aligned). Kept raw so
// that nothing consumes them before phase A: the loads of tile t+1 are issued before phase B of tile t and arrive while
// it runs. YF: the outcome is read as the float the caller passed (wbart; (double) of it is the reference's y), else
// the double latent (probit). AFT: the running fit is part of the tile (streamed from HBM); when it is resident in
// LDS, phase A reads it with its gathers.
template <int R, bool YF = false, bool AFT = true>
struct TileIn {
  static constexpr int S = (16 * R + 63) / 64;  // row slots per thread
  uint32_t la[S], lb[S], qa[S], qb[S];          // leaf id of tree j-1, of tree j; rank on v_prev, on v_cur (a byte a row)
  f4v w[S];
  typename std::conditional<YF, f4v, D4>::type y[S];
  D4 af[AFT ? S : 1];
};

// The first row of the thread's quad in slot s of the tile starting at row slot ii0: tq = 16 t + 4 dw. Addresses are
// clamped to qlast, the wave's last quad that starts below n, so that every load stays inside the chain's columns
// (padded to a multiple of 16 rows); a clamped quad's values are not used.
template <int R, bool RES, bool YF, bool AFT>
__device__ __forceinline__ void tile_load(TileIn<R, YF, AFT>& T, const RowPtrs& P, const uint8_t* __restrict__ la,
                                          const uint8_t* __restrict__ lb, const uint8_t* __restrict__ qa,
                                          const uint8_t* __restrict__ qb, int ii0, uint32_t tq, uint32_t qlast) {
  constexpr int S = TileIn<R, YF, AFT>::S;
#pragma unroll
  for (int s = 0; s < S; s++) {
    const uint32_t r0 = min((uint32_t)((ii0 + 64 * s) * kLanes) + tq, qlast);
    T.la[s] = ldo(reinterpret_cast<const uint32_t*>(la), r0 >> 2);
    T.lb[s] = ldo(reinterpret_cast<const uint32_t*>(lb), r0 >> 2);
    T.qa[s] = ldo(reinterpret_cast<const uint32_t*>(qa), r0 >> 2);
    T.qb[s] = ldo(reinterpret_cast<const uint32_t*>(qb), r0 >> 2);
    T.w[s] = ldo(reinterpret_cast<const f4v*>(P.W), r0 >> 2);
    if constexpr (YF) {
      T.y[s] = ldo(reinterpret_cast<const f4v*>(P.yf), r0 >> 2);
    } else {
      T.y[s].a = ldo(reinterpret_cast<const d2v*>(P.Y), r0 >> 1);
      T.y[s].b = ldo(reinterpret_cast<const d2v*>(P.Y), (r0 >> 1) + 1);
    }
    if constexpr (AFT) {
      T.af[s].a = ldo(reinterpret_cast<const d2v*>(P.AF), r0 >> 1);
      T.af[s].b = ldo(reinterpret_cast<const d2v*>(P.AF), (r0 >> 1) + 1);
    }
  }
}

// the wave's last quad of rows that starts below n (its first quad when the chain has no row in the wave's lanes)
__device__ __forceinline__ uint32_t last_quad(int n, int dw) {
  const int k0 = kLanesPerWave * dw;
  return n > k0 ? (uint32_t)(((n - 1 - k0) & ~(kLanes - 1)) + k0) : (uint32_t)k0;
}

// D: the dense build (two chains per CU, 128 VGPRs), which gathers one row at a time in phase A. RSS: the closing pass
// (residual sum of squares and the posterior-mean sum instead of leaf sums).
template <int R, bool RES, bool FX, bool D = false, bool RSS = false, bool YF, bool AFT>
__device__ void data_pass(Ctl& c, unsigned char* scratch, const RowPtrs& P, int dw, int lane, int jp, int j, int cb,
                          bool apply, bool stats, bool keep, TileIn<R, YF, AFT>& T, bool first_loaded) {
  constexpr bool rss = RSS;
  static_assert(AFT || RES, "a running fit outside LDS is part of the tile");
  using G = Geo<R>;
  const int n = P.n;
  // every control field read here is the same for the whole workgroup: read once into scalar registers
  const int pmode = apply ? (int)ufl(c.pmode) : 0, pnx = ufl(c.pnx), pc = ufl(c.pc), pL = ufl(c.pL), pR = ufl(c.pR),
            pl = ufl(c.pl), pr = ufl(c.pr), pv = ufl(c.pv);
  const int cmode = stats ? (int)ufl(c.cmode[cb]) : 0, cnx = ufl(c.cnx[cb]), cc = ufl(c.cc[cb]), cv = ufl(c.cv[cb]);
  const int nleaf = ufl(c.nleaf[cb]);
  const int ntask = rss ? 2 : 2 * (nleaf + (cmode == 1 ? 2 : 0));
  const bool oh = !D && !FX && !rss && nleaf <= kOneHotLeaves;  // the tree's leaf codes are single bits (leaf_code)
  const size_t ns = (size_t)P.ns;
  uint8_t* __restrict__ lprev = apply ? P.LEAF + (size_t)jp * ns : nullptr;
  const uint8_t* __restrict__ lcur = stats ? P.LEAF + (size_t)j * ns : nullptr;
  const uint8_t* __restrict__ qp = (apply && pmode == 1) ? P.Q + (size_t)pv * ns : nullptr;
  const uint8_t* __restrict__ qc = (stats && cmode == 1) ? P.Q + (size_t)cv * ns : nullptr;
  // unused streams read a valid dummy column so that every load is unconditional
  const uint8_t* __restrict__ la = lprev ? lprev : P.LEAF;
  const uint8_t* __restrict__ lb = lcur ? lcur : P.LEAF;
  const uint8_t* __restrict__ qa = qp ? qp : P.Q;
  const uint8_t* __restrict__ qb = qc ? qc : P.Q;
  const ThR* __restrict__ thr = c.thr[cb];

  // phase-A coordinates: the thread's row quad (16 lane + 4 dw: the wave's four lanes in a row slot); phase-B
  // coordinates: local lane bl, task slot
  const uint32_t tq = opaque((uint32_t)(kLanes * lane + kLanesPerWave * dw));
  const uint32_t qlast = last_quad(n, dw);
  const int bl = lane >> 4, slot = lane & 15, kb = kLanesPerWave * dw + bl;
  double* v0 = reinterpret_cast<double*>(scratch);
  double* v1 = v0 + kLanes * G::kStride + G::kV1Pad;
  uint8_t* tg = reinterpret_cast<uint8_t*>(v1 + kLanes * G::kStride);
  double* v0a = v0 + kLanesPerWave * dw * G::kStride;  // lane 4 dw + u of the quad at + u kStride
  double* v1a = v1 + kLanesPerWave * dw * G::kStride;
  uint8_t* tga = tg + kLanesPerWave * dw * G::kTagStride;
  const double* v0b = v0 + kb * G::kStride;
  const double* v1b = v1 + kb * G::kStride;
  const uint8_t* tgb = tg + kb * G::kTagStride;

  double acc[kGroups];
#pragma unroll
  for (int g = 0; g < kGroups; g++) acc[g] = 0;
  const int nper_kb = kb < n ? (n - kb + kLanes - 1) / kLanes : 0;
  // FX64 sums are order-free: when the tasks take at most half of a lane's 16 slots, the spare slots run replicas of
  // the tasks -- with tp = ntask rounded up to a power of two, slot = task + tp rep for 16 / tp replicas -- each over a
  // contiguous share of the lane's rows of every tile, and the replicas' totals are added at the end of the pass.
  // REF16 keeps one thread per task, walking every row in order.
  const int lgtp = ntask <= 2 ? 1 : (ntask <= 4 ? 2 : (ntask <= 8 ? 3 : 4));
  const bool rp = FX && lgtp < 4;
  const int lgrep = rp ? 4 - lgtp : 0, nrep = 1 << lgrep;
  const int rtask = rp ? (slot & ((1 << lgtp) - 1)) : slot, rep = rp ? (slot >> lgtp) : 0;
  // rows of the birth proposal's halves in this wave's lanes (wave-uniform: counted with ballots; lane 0 takes part in
  // every ballot, as a quad of a higher lane starts below n only if lane 0's does)
  int my_nl = 0, my_nr = 0;

  if (!first_loaded) tile_load<R, RES>(T, P, la, lb, qa, qb, 0, tq, qlast);
  for (int tb = 0; tb < n; tb += G::kTile) {
    const int ii0 = tb / kLanes;
    // ---- phase A: the arithmetic on the tile's loads (issued before the previous tile's phase B). First the quad's
    // leaf ids after tree j-1's change (the move type is the same for the whole workgroup), then per batch of RC rows
    // every LDS gather (leaf values, ranks, the resident fit), all in flight together, then the arithmetic and the
    // stores. Rows of the quad at or past n compute on whatever they load and store into padding that nothing reads
    // as a row: the global pools' padding rows, and in LDS their own slots (the running fit in LDS has a slot for
    // every row of the chain's last 16-row group); a quad past n, or a slot past the tile, does nothing.
    constexpr int S = TileIn<R, YF, AFT>::S;
    constexpr int RC = D ? 1 : 2;  // rows per gather batch (batches keep the registers in budget)
#pragma unroll
    for (int s = 0; s < S; s++) {
      const int idx = lane + 64 * s;  // the slot within the tile
      const uint32_t r0 = (uint32_t)((ii0 + 64 * s) * kLanes) + tq;
      if (idx < G::kTW4 && r0 < (uint32_t)n) {
        // the quad's rows r0 .. r0 + 3 lie in one 16-row group: in the LDS running fit they are afl(r0) + u
        const uint32_t a0 = afl(r0);
        // the quad's leaf ids after the change, a byte a row: a birth splits leaf pnx on the rank, a death merges
        // pl and pr into pnx
        uint32_t lnew = T.la[s];
        if (apply && pmode == 1) {
          lnew = 0;
#pragma unroll
          for (int u = 0; u < kLanesPerWave; u++) {
            const uint32_t lf = (T.la[s] >> (8 * u)) & 255u, qv = (T.qa[s] >> (8 * u)) & 255u;
            const uint32_t lf2 = lf == (uint32_t)pnx ? (qv <= (uint32_t)pc ? (uint32_t)pL : (uint32_t)pR) : lf;
            lnew |= lf2 << (8 * u);
          }
        } else if (apply && pmode == 2) {
          lnew = 0;
#pragma unroll
          for (int u = 0; u < kLanesPerWave; u++) {
            const uint32_t lf = (T.la[s] >> (8 * u)) & 255u;
            const uint32_t lf2 = (lf == (uint32_t)pl || lf == (uint32_t)pr) ? (uint32_t)pnx : lf;
            lnew |= lf2 << (8 * u);
          }
        }
#pragma unroll
        for (int u0 = 0; u0 < kLanesPerWave; u0 += RC) {
          double tho_r[RC], thn_r[RC], thc_r[RC], af_r[RC];
          uint32_t rank_r[RC];
#pragma unroll
          for (int k = 0; k < RC; k++) {
            const int u = u0 + k;
            if (apply) {
              tho_r[k] = c.tho[(T.la[s] >> (8 * u)) & 255u];
              thn_r[k] = c.thn[(lnew >> (8 * u)) & 255u];
            }
            if (!rss) {
              const uint32_t lb_u = (T.lb[s] >> (8 * u)) & 255u;
              const ThR e = thr[lb_u];  // one 16-byte gather
              thc_r[k] = e.th;
              rank_r[k] = e.rank;
            }
            if constexpr (AFT) af_r[k] = d4_at(T.af[s], u);
            else af_r[k] = P.AFl[a0 + (uint32_t)u];
          }
#pragma unroll
          for (int k = 0; k < RC; k++) {
            const int u = u0 + k;
            const uint32_t i = r0 + (uint32_t)u;
            double af = af_r[k];
            if (apply) {
              af = af - tho_r[k] + thn_r[k];
              if (RES) P.AFl[a0 + (uint32_t)u] = af;
              else sto(P.AF, i, af);
            }
            double y;
            if constexpr (YF) y = (double)T.y[s][u];
            else y = d4_at(T.y[s], u);
            const double w = (double)T.w[s][u];
            if (rss) {
              const double e = y - af;
              const double c2 = (w * e) * e;
              v0a[u * G::kStride + idx] = 0.0;
              v1a[u * G::kStride + idx] = FX ? __longlong_as_double(fx_round(c2, 16777216.0)) : c2;
              tga[u * G::kTagStride + idx] = 0;
              if (keep) sto(P.TM, i, ldo((const double*)P.TM, i) + af);
            } else {
              const double rr = y - af + thc_r[k];
              const double wr = w * rr;
              // tag: the leaf's code (leaf_code), bit 6 / 7 for the left / right half of the birth proposal's node
              uint32_t tag = rank_r[k];
              if (cmode == 1) {
                const bool split = ((T.lb[s] >> (8 * u)) & 255u) == (uint32_t)cnx && i < (uint32_t)n;
                const bool left = ((T.qb[s] >> (8 * u)) & 255u) <= (uint32_t)cc;
                const bool isl = split && left, isr = split && !left;
                my_nl += __popcll(__ballot(isl));
                my_nr += __popcll(__ballot(isr));
                tag |= isl ? 64u : (isr ? 128u : 0u);
              }
              v0a[u * G::kStride + idx] = FX ? __longlong_as_double(fx_round(w, 4294967296.0)) : w;
              v1a[u * G::kStride + idx] = FX ? __longlong_as_double(fx_round(wr, 4294967296.0)) : wr;
              tga[u * G::kTagStride + idx] = (uint8_t)tag;
            }
          }
        }
        // the quad's new leaf ids in one dword (bytes past n land in the column's padding, and stay node ids: they
        // start at 0 and only ever receive a leaf id of their own row's old value)
        if (apply) sto(reinterpret_cast<uint32_t*>(lprev), r0 >> 2, lnew);
      }
    }
    wave_sync();
    // ---- phase B: this thread's tasks walk lane kb's rows of the tile in row order. task = 2 rank + kind: kind 0
    // sums w, kind 1 sums w*r; ranks below nleaf are leaves (rank field, or in the one-chain-per-CU REF16 build the
    // rank's bit for trees of at most kOneHotLeaves leaves), then the split halves (side field 1 and 2, bits 6 and
    // 7); the closing pass sums every row (an empty field matches bit 0). A slot past ntask sums nothing.
    // the next tile's loads go out now and arrive during phase B
    if (tb + G::kTile < n) tile_load<R, RES>(T, P, la, lb, qa, qb, ii0 + G::kTW4, tq, qlast);
    const int cnt = min(G::kTW4, max(0, nper_kb - ii0));
    // this thread's rows of the lane's tile: all of them, or a replica's share (whole 16-row blocks but the last)
    int r0 = 0, rn = cnt;
    if (rp) {
      const int share = (((cnt + nrep - 1) >> lgrep) + 15) & ~15;
      r0 = min(cnt, rep * share);
      rn = min(cnt - r0, share);
    }
#pragma unroll
    for (int g = 0; g < kGroups; g++) {
      const int task = rtask + kTaskSlots * g;
      if (task < ntask && (!rp || g == 0)) {
        const int q = task >> 1;
        const double* src = ((task & 1) ? v1b : v0b) + r0;
        if (!D && !FX && oh) {
          const uint32_t bit = q < nleaf ? (uint32_t)q : (q == nleaf ? 6u : 7u);
          acc[g] = lane_sum<FX, !D && !FX>(tgb + r0, src, rn, bit, 0u, 0u, acc[g]);
        } else {
          const uint32_t sh = (!rss && q >= nleaf) ? 6u : 0u;
          const uint32_t wd = rss ? 0u : (q < nleaf ? 5u : 2u);
          const uint32_t pat = rss ? 1u : (q < nleaf ? (1u << q) : (q == nleaf ? 2u : 4u));
          acc[g] = lane_sum<FX, false>(tgb + r0, src, rn, sh, wd, pat, acc[g]);
        }
      }
    }
    wave_sync();
  }
  if (lane == 0 && my_nl) atomicAdd(&c.nL, my_nl);
  if (lane == 0 && my_nr) atomicAdd(&c.nR, my_nr);
  if (rp) {
    // the first replica adds the others' totals (int64 adds: exact in any order); replica k of a task sits tp k slots
    // above it, in the same 16-slot group of the lane
    const long long a = __double_as_longlong(acc[0]);
    long long t = a;
    for (int k = 1; k < nrep; k++) t += __shfl(a, lane + (k << lgtp));
    if (rep == 0 && rtask < ntask) c.part[rtask][kb] = __longlong_as_double(t);
  } else {
#pragma unroll
    for (int g = 0; g < kGroups; g++) {
      const int task = slot + kTaskSlots * g;
      if (task < ntask) c.part[task][kb] = acc[g];
    }
  }
}

// The next pass's first tile, loaded between the end of pass j and the end of tree j's Metropolis-Hastings step: the
// columns it reads are fixed by then (tree j's leaf ids, tree j+1's, and the ranks on both proposal variables); whether
// tree j's birth was accepted only decides whether phase A uses the ranks on its variable. After the last tree it is
// the closing pass's tile.
template <int R, bool RES, bool YF, bool AFT>
__device__ __forceinline__ void load_next_first_tile(TileIn<R, YF, AFT>& T, const Ctl& c, const RowPtrs& P, int dw,
                                                     int lane,
                                                     int j, int m, int cb) {
  const int n = P.n;
  const uint32_t tq = opaque((uint32_t)(kLanes * lane + kLanesPerWave * dw));
  const bool last = j + 1 == m;
  const size_t ns = (size_t)P.ns;
  const uint8_t* la = P.LEAF + (size_t)j * ns;
  const uint8_t* lb = last ? P.LEAF : P.LEAF + (size_t)(j + 1) * ns;
  const uint8_t* qa = ufl(c.cmode[cb]) == 1 ? P.Q + (size_t)ufl(c.cv[cb]) * ns : P.Q;
  const uint8_t* qb = (!last && ufl(c.cmode[cb ^ 1]) == 1) ? P.Q + (size_t)ufl(c.cv[cb ^ 1]) * ns : P.Q;
  tile_load<R, RES>(T, P, la, lb, qa, qb, 0, tq, last_quad(n, dw));
}

// ------------------------------------------------------------------ probit latent redraw (all waves)
// The control wave generates attempt slots into the LDS ring, running kAhead slots ahead of the current window; the
// data waves decide the window's attempts and, meanwhile, evaluate and store the previous window's accepted latents;
// then the control wave scans the masks. Starts from the generator state in c.pst and leaves there the state after
// the last row's accepted attempt. Control flow is uniform over the workgroup.
__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, int l) {
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ double readlane_d(double v, int l) {
  return __longlong_as_double((long long)readlane_u64((uint64_t)__double_as_longlong(v), l));
}

// 64 slots from F (a multiple of 64): lane k receives slot F + k's two state words and, every kCkpt slots, a checkpoint
__device__ __forceinline__ void ring_batch(ProbitRing& B, Rng& h, uint32_t& F, int lane, uint32_t k11, uint32_t k21) {
  uint32_t r1 = 0, r2 = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
#pragma unroll 8
  for (int k = 0; k < 64; k++) {
    if (k % kCkpt == 0) {
      const bool mine = lane == k / kCkpt;
      c0 = mine ? h.s0 : c0;
      c1 = mine ? h.s1 : c1;
      c2 = mine ? h.s2 : c2;
      c3 = mine ? h.s3 : c3;
    }
    const bool mine = lane == k;
    r1 = mine ? h.s1 : r1;
    rng_step(h, k11, k21);
    r2 = mine ? h.s1 : r2;
    rng_step(h, k11, k21);
  }
  const uint32_t base = F % kRing;
  *reinterpret_cast<uint2*>(B.raw[base + lane]) = make_uint2(r1, r2);
  if (lane < 64 / kCkpt) *reinterpret_cast<uint4*>(B.ck[base / kCkpt + lane]) = make_uint4(c0, c1, c2, c3);
  F += 64;
}
__device__ __forceinline__ void ring_generate(ProbitRing& B, Rng& h, uint32_t& F, uint32_t upto, int lane,
                                              uint32_t k11, uint32_t k21) {
  while (F < upto) ring_batch(B, h, F, lane, k11, k21);
}
__device__ __forceinline__ uint32_t slot_x1(const ProbitRing& B, uint32_t slot) { return starstar(B.raw[slot % kRing][0]); }
__device__ __forceinline__ uint32_t slot_x2(const ProbitRing& B, uint32_t slot) { return starstar(B.raw[slot % kRing][1]); }

// Not inlined: it runs once per sweep, and with its own register budget it does not spill inside the scans.
template <bool RES>
__device__ __attribute__((noinline)) void probit_latents(int n_, double off_, const double* AFg_, int af_off_,
                                                         const uint8_t* YB_, double* Y_, LibmTables tb_, uint32_t g0,
                                                         uint32_t g1, uint32_t g2, uint32_t g3, uint32_t k11_,
                                                         uint32_t k21_) {
  // the arguments are the same in every lane: say so, so that loop control and the generator stay scalar
  const int n = (int)ufl((uint32_t)n_);
  const double off = ufl_t(off_);
  const int af_off = (int)ufl((uint32_t)af_off_);
  const uint8_t* __restrict__ YB = ufl_t(YB_);
  double* __restrict__ Y = ufl_t(Y_);
  const LibmTables tb{ufl_t(tb_.logf_t), ufl_t(tb_.cosf_t), ufl_t(tb_.logd_t)};
  const uint32_t k11 = ufl(k11_), k21 = ufl(k21_);
  ProbitRing& B = *reinterpret_cast<ProbitRing*>(fbh_dyn);
  const double* __restrict__ AF = RES ? reinterpret_cast<const double*>(fbh_dyn + af_off) : ufl_t(AFg_);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  Rng h{0, 0, 0, 0};
  uint32_t F = 0;  // slots generated so far (control wave)
  if (wave == 0) {
    h = Rng{ufl(g0), ufl(g1), ufl(g2), ufl(g3)};
    ring_generate(B, h, F, kAhead, lane, k11, k21);
  }
  __syncthreads();
  int row0 = 0, prow0 = 0, pdone = 0;
  int rho16 = 8;  // predicted rejections per row, in 1/16, from the windows so far
  uint32_t S = 0, pS = 0;
  for (;;) {
    const int reff = min(kPR, n - row0);
    if (wave == 0) {
      if (reff > 0) ring_generate(B, h, F, S + kAhead, lane, k11, k21);
    } else {
      // this window's rows (lane k holds row dw + 4k) and the previous window's accepted attempts: row loads first,
      // then the image gathers of the accepted attempts, whose latency the decisions cover
      const int dw = wave - 1;
      const int lr = dw + kDataWaves * lane;
      double fv = 0.;
      int ybv = 0;
      if (reff > 0 && lr < reff) {
        fv = RES ? AF[afl(row0 + lr)] : gld(AF + row0 + lr);
        ybv = gld(YB + row0 + lr);
      }
      const int j = tid - 64;
      const bool fin = j < pdone;
      double fj = 0.;
      int ybj = 0;
      uint32_t x1 = 0, x2 = 0;
      float lfj = 0.f, cfj = 0.f;
      double ldj = 0.;
      if (fin) {
        fj = RES ? AF[afl(prow0 + j)] : gld(AF + prow0 + j);
        ybj = gld(YB + prow0 + j);
        const uint32_t k = pS + (uint32_t)j + B.acc[j];
        x1 = slot_x1(B, k);
        x2 = slot_x2(B, k);
        lfj = gld(tb.logf_t + (x1 >> 8));
        cfj = gld(tb.cosf_t + (x2 >> 8));
        ldj = gld(tb.logd_t + (x1 >> 8));
      }
      // decisions: data wave dw takes rows dw + 4i (i < 16), four rows per step; lane 16g + b decides offset
      // band_start(l) + b of row l = dw + 4 (4k + g)
      if (reff > 0) {
        // each lane's own row's truncation point and rate once per window (an FP64 square root), handed to the
        // sixteen lanes that decide that row's offsets
        const LatentRow mine = latent_row(fv, ybv, off);
        for (int k = 0; k < kPR / (4 * kDataWaves); k++) {
          const int i = 4 * k + (lane >> 4);
          const int l = dw + kDataWaves * i;
          LatentRow r;
          r.tau = __shfl(mine.tau, i);
          r.lam = __shfl(mine.lam, i);
          r.tauf = __shfl(mine.tauf, i);
          r.lamf = __shfl(mine.lamf, i);
          r.mean = 0.;  // not used by the decision
          r.pos = false;
          bool ok = false;
          if (l < reff) {
            const uint32_t slot = S + (uint32_t)(l + band_start(l, rho16) + (lane & 15));
            ok = latent_accepts(r, slot_x1(B, slot), slot_x2(B, slot), tb);
          }
          const uint64_t msk = __ballot(ok);
          if (lane < 4) {
            const int lg = dw + kDataWaves * (4 * k + lane);
            if (lg < reff) B.dmask[lg] = (uint32_t)(msk >> (16 * lane)) & 0xFFFFu;
          }
        }
      }
      if (fin) {
        const LatentRow rj = latent_row(fj, ybj, off);
        double z;
        if (rj.tau <= 0.) {
          z = (double)(__builtin_sqrtf(-2.f * lfj) * cfj);
        } else {
          z = -ldj / rj.lam + rj.tau;
        }
        const double v = z + rj.mean;
        gst(Y + prow0 + j, rj.pos ? v : -v);
      }
    }
    __syncthreads();
    if (reff <= 0) break;
    if (wave == 0) {
      // row l starts at offset r (the rejections so far) and accepts at the first set bit at or above it. A step
      // resolves at once the run of rows from `done` that accept at the current offset (one ballot: their bit at r is
      // set), then the first row that does not: its next set bit above r, which advances r. Steps = rows that reject
      // at least once, not rows. Lane l holds its row's band start and mask; pk packs both for the readlane of row f.
      const uint32_t bsl = lane < reff ? (uint32_t)band_start(lane, rho16) : 0u;
      const uint32_t mkl = lane < reff ? B.dmask[lane] : 0u;
      const uint32_t pk = (bsl << 16) | mkl;
      const uint64_t below_reff = reff >= 64 ? ~0ull : ((1ull << reff) - 1ull);
      int r = 0, done = 0;
      uint32_t accv = 0;
#pragma unroll 1
      while (done < reff) {
        const uint32_t sh = (uint32_t)r - bsl;
        const bool ok = sh < (uint32_t)kBand && ((mkl >> sh) & 1u);
        const uint64_t notok = below_reff & (~0ull << done) & ~(uint64_t)__ballot(ok);
        const int f = notok ? (int)__builtin_ctzll(notok) : reff;
        accv = (lane >= done && lane < f) ? (uint32_t)r : accv;
        done = f;
        if (done >= reff) break;
        const uint32_t x = __builtin_amdgcn_readlane(pk, f);
        const uint32_t shf = (uint32_t)r - (x >> 16);
        const uint32_t mk = shf < (uint32_t)kBand ? (x & 0xFFFFu) >> shf : 0u;
        if (!mk) break;
        r += __builtin_ctz(mk);
        accv = lane == f ? (uint32_t)r : accv;
        done = f + 1;
      }
      if (lane < done) B.acc[lane] = (uint8_t)accv;
      int used = done + r;
      if (done == 0) {
        // the window's first row rejected kPT times in a row: its attempts one by one, generating as needed
        const LatentRow r0 = latent_row(RES ? AF[afl(row0)] : gld(AF + row0), (int)gld(YB + row0), off);
        uint32_t k = 0;
        for (;; k++) {
          if (S + k >= F) ring_generate(B, h, F, S + k + 1, lane, k11, k21);
          if (latent_accepts_exact(r0, slot_x1(B, S + k), slot_x2(B, S + k), tb)) break;
        }
        const double z = latent_value(r0, slot_x1(B, S + k), slot_x2(B, S + k), tb);
        if (lane == 0) gst(Y + row0, z);
        used = (int)k + 1;
      }
      // keep the next window's slots generated
      ring_generate(B, h, F, S + (uint32_t)used + kNeed, lane, k11, k21);
      if (lane == 0) {
        B.p_done = done;
        B.p_used = used;
      }
    }
    __syncthreads();
    const int done = (int)ufl((uint32_t)B.p_done);
    const int used = (int)ufl((uint32_t)B.p_used);
    if (done >= 8) rho16 = min(kRhoMax, (3 * rho16 + (16 * (used - done)) / done + 2) / 4);
    prow0 = row0;
    pS = S;
    pdone = done;
    row0 += done > 0 ? done : 1;
    S += (uint32_t)used;
  }
  if (wave == 0) {
    // the state before slot S: its checkpoint, stepped over the slots since
    const uint32_t q = (S / kCkpt) % (kRing / kCkpt);
    Rng e{ufl(B.ck[q][0]), ufl(B.ck[q][1]), ufl(B.ck[q][2]), ufl(B.ck[q][3])};
    for (uint32_t i = 0; i < 2 * (S % kCkpt); i++) rng_step(e, k11, k21);
    if (lane == 0) {
      B.fin[0] = e.s0;
      B.fin[1] = e.s1;
      B.fin[2] = e.s2;
      B.fin[3] = e.s3;
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------ row split: pass totals across the workgroups
// Every member of a split chain runs the same control steps on the same generator; after each pass the members add
// their fixed-point totals (int64, order-free) so that every member continues from the same sums. Protocol, all on
// device-scope (agent) atomics so that members on different XCDs see each other's values at the coherence point:
// store this member's totals, wait until the stores have completed, count the arrival, wait for all members'
// arrivals, read the totals of every member. Slots alternate between two buffers (a member can reuse a buffer only
// after every member has arrived at the next exchange, i.e. has finished reading it). A wait that exceeds the spin
// limit (members that are not resident together) sets xerr and stops waiting in this and later exchanges.
constexpr int kXw = kMaxTasks + 2;    // task totals, then the two split counts
constexpr int kSpinLimit = 1 << 20;   // about a second of polling
__device__ void row_split_exchange(Ctl& c, const KArgs& a, int cb, bool rss, uint32_t& xseq, int& xerr, int lane) {
  const int W = c.cd.split, me = c.cd.member, grp = c.cd.group;
  const int ntask = rss ? 2 : 2 * (c.nleaf[cb] + (c.cmode[cb] == 1 ? 2 : 0));
  unsigned long long* X = a.xbuf + ((size_t)grp * 2 + (xseq & 1u)) * (size_t)W * kXw;
  for (int t = lane; t < kXw; t += 64) {
    long long v = 0;
    if (t < ntask)
      for (int k = 0; k < kLanes; k++) v += __double_as_longlong(c.part[t][k]);
    else if (t == kXw - 2)
      v = c.nL;
    else if (t == kXw - 1)
      v = c.nR;
    __hip_atomic_store(X + (size_t)me * kXw + t, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed before its arrival is counted
  if (lane == 0) {
    __hip_atomic_fetch_add(a.xctr + grp, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t target = (xseq + 1u) * (uint32_t)W;
    for (int spin = 0; !xerr; spin++) {
      if (__hip_atomic_load(a.xctr + grp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      if (spin >= kSpinLimit) xerr = 1;
      __builtin_amdgcn_s_sleep(2);
    }
  }
  xerr = (int)ufl((uint32_t)__shfl(xerr, 0));
  for (int t = lane; t < kXw; t += 64) {
    long long v = 0;
    for (int k = 0; k < W; k++)
      v += (long long)__hip_atomic_load(X + (size_t)k * kXw + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (t < ntask) {
      c.part[t][0] = __longlong_as_double(v);
      for (int k = 1; k < kLanes; k++) c.part[t][k] = 0.0;
    } else if (t == kXw - 2) {
      c.nL = (int)v;
    } else if (t == kXw - 1) {
      c.nR = (int)v;
    }
  }
  xseq++;
  wave_sync();
}

// ------------------------------------------------------------------ the sweep kernel
// Residency: the dispatcher reserves registers for a 5-wave workgroup as if each SIMD held two of its waves, so a
// second workgroup fits on a CU only if four waves fit per SIMD, i.e. at most 128 VGPRs (with 168, 512 workgroups on
// 256 CUs were measured to run as two consecutive rounds of 256). D = 1, the dense build: 128 VGPRs, two chains per
// CU, for batches with more workgroups than CUs. D = 0: 168 VGPRs (three waves per SIMD), one chain per CU, fewer
// spills. PB: probit (latent redraw after each sweep, sigma fixed) -- a separate instantiation so that the wbart
// kernel carries none of its code.
template <int R, bool RES, bool FX, bool PB, bool D>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(D ? 4 : 3))) void k_sweeps(
    KArgs a) {
  __shared__ Ctl c;
  unsigned char* const dyn = fbh_dyn;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: the role branches below are scalar
  const bool ctl = wave == 0;
  const int dw = wave - 1;
  const int chain = a.order[blockIdx.x];  // the packing: which chain this workgroup runs
  if (ctl) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.desc + chain);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&c.cd);
    for (int w = lane; w < (int)(sizeof(ChainDesc) / 4); w += 64) dst[w] = src[w];
    c.PG[lane] = a.PG[lane];
    if (lane < 2) c.rot[lane] = lane ? 21u : 11u;
  }
  __syncthreads();
  const int n = (int)ufl((uint32_t)c.cd.n), m = a.m, cap = a.node_cap;
  const uint64_t row_off = ufl_t(c.cd.row_off);
  RowPtrs P;
  P.AF = a.AF + row_off;
  P.AFl = reinterpret_cast<double*>(dyn + a.af_off);
  P.TM = a.TM + row_off;
  P.Y = a.Y + row_off;
  P.yf = a.yf + row_off;
  P.W = a.W + row_off;
  P.Q = a.q + ufl_t(c.cd.q_off);
  P.LEAF = a.leaf + ufl_t(c.cd.leaf_off);
  P.n = n;
  P.probit = a.probit;
  P.ns = ufl(c.cd.ns);
  TreeDev* TR = a.trees + c.cd.tree_off;
  ChainState* cs = a.state + chain;
  const LibmTables tb = a.tb;
  const double t2 = (double)c.cd.tau * c.cd.tau;
  const double a_prior = 1 / t2;
  if (RES && !ctl)
    for (int i = tid - 64; i < n; i += kBlock - 64) P.AFl[afl(i)] = P.AF[i];

  // The two roles run separate loops with the same sequence of workgroup barriers per sweep (1 + 2 m + 1, and 2 more
  // for the probit redraw), so that neither role's live registers are held across the other's code.
  if (ctl) {
    Rng g{ufl(cs->rng[0]), ufl(cs->rng[1]), ufl(cs->rng[2]), ufl(cs->rng[3])};
    double sigma = cs->sigma;
    int n_birth = cs->n_birth, n_death = cs->n_death, n_cap = cs->n_cap;
    uint32_t xseq = ufl(cs->xseq);
    int xerr = (int)ufl((uint32_t)cs->xerr);
    const bool split = FX && c.cd.split > 1;
    for (int it = a.it0; it < a.it1; it++) {
      const bool keep = it >= a.nskip;
      if (lane == 0) {
        c.inv_s2 = 1.0 / (sigma * sigma);
        c.pmode = 0;
        c.nL = c.nR = 0;  // the data waves add their split counts into these during each pass
      }
      copy_tree(&c.tree[0], &TR[0], lane);
      wave_sync();
      ctl_propose(c, 0, g, cap, lane, !D && !FX);
      __syncthreads();
      __builtin_amdgcn_s_setprio(0);
      int cb = 0;
      for (int j = 0; j < m; j++) {
        // the previous update's tree (buffer cb ^ 1) is stored before that buffer receives tree j+1
        if (j > 0) {
          const TreeDev& tp = c.tree[cb ^ 1];
          if (keep && a.keep_trees && c.cd.keep)
            a.trees_out[c.cd.tout_off + ((size_t)(it - a.nskip) * m + (j - 1)) * kNodes + lane] = pack_word(tp, lane);
          copy_tree(&TR[j - 1], &tp, lane);
        }
        // tree j+1's words are requested first so that their latency overlaps the generator stepping
        const bool nxt = j + 1 < m;
        const uint64_t* tsrc = reinterpret_cast<const uint64_t*>(&TR[nxt ? j + 1 : j]);
        const uint64_t tw0 = tsrc[lane];
        const uint64_t tw1 = tsrc[lane < kTreeWords - 64 ? 64 + lane : lane];
        // this update's generator outputs: the MH uniform (:482), then two per slot for the leaf normals (:521-523)
        const uint32_t x0 = rng_next(g);
        // the 2 cap steps keep only each output's raw s1 word (lane s gets slot s's pair); the output function is then
        // applied on all lanes at once
        const uint32_t k11 = ufl(c.rot[0]), k21 = ufl(c.rot[1]);
        uint32_t r1 = 0, r2 = 0;
        for (int s = 0; s < cap; s++) {
          const uint32_t a1 = g.s1;
          rng_step(g, k11, k21);
          const uint32_t a2 = g.s1;
          rng_step(g, k11, k21);
          r1 = lane == s ? a1 : r1;
          r2 = lane == s ? a2 : r2;
        }
        const uint32_t my1 = starstar(r1), my2 = starstar(r2);
        if (lane == 0) {
          c.rng_mark[0] = g.s0;
          c.rng_mark[1] = g.s1;
          c.rng_mark[2] = g.s2;
          c.rng_mark[3] = g.s3;
        }
        // libm-image gathers for the MH uniform and for the slots that can be leaves after the MH step; consumed after
        // the next proposal
        const bool want = lane < cap && BITM(c.cand[cb], lane);
        const double luv = tb.logd_t[x0 >> 8];
        const float lf = tb.logf_t[want ? (my1 >> 8) : 0u];
        const float cf = tb.cosf_t[want ? (my2 >> 8) : 0u];
        if (nxt) {
          uint64_t* tdst = reinterpret_cast<uint64_t*>(&c.tree[cb ^ 1]);
          tdst[lane] = tw0;
          if (lane < kTreeWords - 64) tdst[64 + lane] = tw1;
          wave_sync();
          ctl_propose(c, cb ^ 1, g, cap, lane, !D && !FX);
        }
        if (want) c.z[lane] = __builtin_sqrtf(-2.f * lf) * cf;
        if (lane == 0) c.lu = luv;
        __syncthreads();
        __builtin_amdgcn_s_setprio(3);
        // ---- lane sums, Metropolis-Hastings (:469-518), leaf values (:520-528)
        if (split) row_split_exchange(c, a, cb, false, xseq, xerr, lane);
        const int cmode = c.cmode[cb];
        if (FX) ctl_combine<true>(c, cb, false, cmode, lane);
        else ctl_combine<false>(c, cb, false, cmode, lane);
        TreeDev& t = c.tree[cb];
        const double inv_s2 = c.inv_s2;
        const double lu = c.lu;
        const double pr = c.prr[cb];
        double lalpha = 0;
        int accepted = 0, capped = 0, nmode = 0, nL_ = 0, nR_ = 0, nl_ = 0, nr_ = 0;
        const int cnx = c.cnx[cb];
        const int nLc = c.nL, nRc = c.nR;
        if (cmode == 1) {
          if (nLc >= 5 && nRc >= 5) {
            const double bl = c.sL0 * inv_s2, Ml = c.sL1 * inv_s2, br = c.sR0 * inv_s2, Mr = c.sR1 * inv_s2;
            // the three leaf likelihoods on lanes 0-2, added in the reference's order
            const double h = lane == 0 ? lh(bl, Ml, t2) : (lane == 1 ? lh(br, Mr, t2) : lh(bl + br, Ml + Mr, t2));
            lalpha = c.lpr[cb] + __shfl(h, 0) + __shfl(h, 1) - __shfl(h, 2);
            if (lalpha > 0) lalpha = 0;
            if (lu < lalpha) {
              const int L = first_free(t.used, -1, cap);
              const int Rr = L < 0 ? -1 : first_free(t.used, L, cap);
              if (L < 0 || Rr < 0) {
                n_cap++;
                capped = 1;
              } else {
                if (lane == 0) {
                  t.v[cnx] = (uint8_t)c.cv[cb];
                  t.c[cnx] = (uint8_t)c.cc[cb];
                  t.l[cnx] = (uint8_t)L;
                  t.r[cnx] = (uint8_t)Rr;
                  t.par[L] = t.par[Rr] = (uint8_t)cnx;
                  t.l[L] = t.r[L] = t.l[Rr] = t.r[Rr] = 0;
                  t.d[L] = t.d[Rr] = (uint8_t)(t.d[cnx] + 1);
                  t.used |= (1ull << L) | (1ull << Rr);
                  t.leaf = (t.leaf & ~(1ull << cnx)) | (1ull << L) | (1ull << Rr);
                  c.S0[L] = c.sL0;
                  c.S1[L] = c.sL1;
                  c.S0[Rr] = c.sR0;
                  c.S1[Rr] = c.sR1;
                }
                wave_sync();
                const int gL = count_goodvars_wave(t, L, c.cd.nc, c.cd.p, lane);
                const int gR = count_goodvars_wave(t, Rr, c.cd.nc, c.cd.p, lane);
                if (lane == 0) {
                  t.ngv[L] = (uint8_t)gL;
                  t.ngv[Rr] = (uint8_t)gR;
                }
                nmode = 1;
                nL_ = L;
                nR_ = Rr;
                n_birth++;
                accepted = 1;
              }
            }
          }
        } else if (cmode == 2) {
          const int l = t.l[cnx], r = t.r[cnx];
          const double bl = c.S0[l] * inv_s2, Ml = c.S1[l] * inv_s2, br = c.S0[r] * inv_s2, Mr = c.S1[r] * inv_s2;
          const double h = lane == 0 ? lh(bl + br, Ml + Mr, t2) : (lane == 1 ? lh(bl, Ml, t2) : lh(br, Mr, t2));
          lalpha = c.lpr[cb] + __shfl(h, 0) - __shfl(h, 1) - __shfl(h, 2);
          if (lalpha > 0) lalpha = 0;
          if (lu < lalpha) {
            const double m0 = c.S0[l] + c.S0[r], m1 = c.S1[l] + c.S1[r];
            wave_sync();
            if (lane == 0) {
              c.S0[cnx] = m0;
              c.S1[cnx] = m1;
              t.used &= ~((1ull << l) | (1ull << r));
              t.leaf = (t.leaf & ~((1ull << l) | (1ull << r))) | (1ull << cnx);
              t.l[cnx] = t.r[cnx] = t.v[cnx] = t.c[cnx] = 0;
            }
            nmode = 2;
            nl_ = l;
            nr_ = r;
            n_death++;
            accepted = 1;
          }
        }
        wave_sync();
        // leaf values: the normals of the used leaves (gathered during the pass)
        const int s = lane;
        c.tho[s] = c.thr[cb][s].th;
        if (s < cap && BITM(t.leaf, s) && BITM(t.used, s)) {
          const double b = c.S0[s] * inv_s2, M = c.S1[s] * inv_s2;
          t.th[s] = (float)(M / (a_prior + b) + (double)c.z[s] / __builtin_sqrt(a_prior + b));
        }
        wave_sync();
        c.thn[s] = (double)t.th[s];
        // tree j goes back to global memory (and to the kept draws) during the next pass, off the critical path
        if (lane == 0) {
          c.pmode = nmode;
          c.pnx = cnx;
          c.pv = c.cv[cb];
          c.pc = c.cc[cb];
          c.pL = nL_;
          c.pR = nR_;
          c.pl = nl_;
          c.pr = nr_;
          c.nL = c.nR = 0;
          if (a.trace) {
            const long idx = (long)(it - a.it0) * m + j;
            if (idx < a.trace_cap) {
              TraceRec& tr = a.trace[(long)chain * a.trace_cap + idx];
              tr.it = it;
              tr.j = j;
              tr.mode = cmode;
              tr.nx = cnx;
              tr.v = c.cv[cb];
              tr.c = c.cc[cb];
              tr.nl = nLc;
              tr.nr = nRc;
              tr.accepted = accepted;
              tr.ngood = c.ngood[cb];
              tr.nnogs = c.nnogs[cb];
              tr.cap = capped;
              tr.pr = pr;
              tr.lalpha = lalpha;
              tr.lu = lu;
              tr.sl0 = c.sL0;
              tr.sl1 = c.sL1;
              tr.sr0 = c.sR0;
              tr.sr1 = c.sR1;
              for (int w = 0; w < 4; w++) tr.rng_after[w] = c.rng_mark[w];
            }
          }
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(0);
        cb ^= 1;
      }
      // closing pass (:536-541) on the data waves; meanwhile the last tree (buffer cb ^ 1) is stored
      const TreeDev& tp = c.tree[cb ^ 1];
      if (keep && a.keep_trees && c.cd.keep)
        a.trees_out[c.cd.tout_off + ((size_t)(it - a.nskip) * m + (m - 1)) * kNodes + lane] = pack_word(tp, lane);
      copy_tree(&TR[m - 1], &tp, lane);
      __syncthreads();
      // from here to the first pass of the next sweep (sigma, the first proposal) the control wave is on the critical
      // path again; in the probit redraw the data waves' decisions are, and it stays below them (measured faster)
      if (!PB) __builtin_amdgcn_s_setprio(3);
      if (PB) {
        // the latents are redrawn by every wave (:543-547); sigma stays at 1
        if (lane == 0) {
          c.pst_[0] = g.s0;
          c.pst_[1] = g.s1;
          c.pst_[2] = g.s2;
          c.pst_[3] = g.s3;
        }
      } else {
        if (split) row_split_exchange(c, a, cb, true, xseq, xerr, lane);
        if (FX) ctl_combine<true>(c, cb, true, 0, lane);
        else ctl_combine<false>(c, cb, true, 0, lane);
        const float nul = c.cd.nu * c.cd.lambda;
        const float nnu = (float)c.cd.n_total + c.cd.nu;
        sigma = __builtin_sqrt(((double)nul + c.rss) / (2.0 * rgamma_dev(g, (double)nnu / 2.0, tb)));
      }
      if (lane == 0) a.sigma_out[c.cd.sig_off + it] = (float)sigma;
      if (PB && !a.pr_ext) {
        __syncthreads();
        probit_latents<RES>(n, (double)c.cd.offset, P.AF, a.af_off, a.ybin + c.cd.row_off, a.Y + c.cd.row_off, tb,
                            ufl(c.pst_[0]), ufl(c.pst_[1]), ufl(c.pst_[2]), ufl(c.pst_[3]), ufl(c.rot[0]),
                            ufl(c.rot[1]));
        const ProbitRing& B = *reinterpret_cast<const ProbitRing*>(dyn);
        g = Rng{ufl(B.fin[0]), ufl(B.fin[1]), ufl(B.fin[2]), ufl(B.fin[3])};
      }
      __syncthreads();
    }
    if (lane == 0) {
      cs->rng[0] = g.s0;
      cs->rng[1] = g.s1;
      cs->rng[2] = g.s2;
      cs->rng[3] = g.s3;
      cs->sigma = sigma;
      cs->it = a.it1;
      cs->n_birth = n_birth;
      cs->n_death = n_death;
      cs->n_cap = n_cap;
      cs->xseq = xseq;
      cs->xerr = xerr;
    }
  } else {
    TileIn<R, !PB, !RES> T;
    // Wave priorities (they only order issue among the waves sharing a SIMD): the data waves run at 2; the control
    // wave at 0 while it prepares the next update during the pass (it finishes with thousands of cycles to spare, and
    // the data wave beside it on its SIMD sets the pass's length) and at 3 from the pass barrier to the end of the
    // update, where its lane sums, Metropolis-Hastings step and leaf values are the chain's critical path.
    __builtin_amdgcn_s_setprio(2);
    for (int it = a.it0; it < a.it1; it++) {
      const bool keep = it >= a.nskip;
      __syncthreads();
      int cb = 0;
      bool pre = false;
      for (int j = 0; j < m; j++) {
        data_pass<R, RES, FX, D, false>(c, dyn, P, dw, lane, j - 1, j, cb, j > 0, true, false, T, pre);
        __syncthreads();
        // while the control wave decides tree j, the next pass's first tile is loaded
        load_next_first_tile<R, RES>(T, c, P, dw, lane, j, m, cb);
        pre = true;
        __syncthreads();
        cb ^= 1;
      }
      data_pass<R, RES, FX, D, true>(c, dyn, P, dw, lane, m - 1, m, cb, true, false, keep, T, pre);
      __syncthreads();
      if (PB && !a.pr_ext) {
        __syncthreads();
        probit_latents<RES>(n, (double)c.cd.offset, P.AF, a.af_off, a.ybin + c.cd.row_off, a.Y + c.cd.row_off, tb,
                            ufl(c.pst_[0]), ufl(c.pst_[1]), ufl(c.pst_[2]), ufl(c.pst_[3]), ufl(c.rot[0]),
                            ufl(c.rot[1]));
      }
      __syncthreads();
    }
    if (RES)
      for (int i = tid - 64; i < n; i += kBlock - 64) P.AF[i] = P.AFl[afl(i)];
  }
}

// Chain initialisation: root-only trees, generator seed + 16 discarded outputs, the initial probit latents
// (src/bart_kernel.cpp:367-386) -- those by one thread here, or (pr_ext) by the k_pr_* kernels after this one, as
// the redraw of a zero running fit.
__global__ __launch_bounds__(256) void k_init(KArgs a, const unsigned* seeds) {
  const int tid = threadIdx.x;
  const int chain = blockIdx.x;
  const ChainDesc cd = a.desc[chain];
  TreeDev* TR = a.trees + cd.tree_off;
  for (int w = tid; w < a.m * (int)(sizeof(TreeDev) / 8); w += 256) reinterpret_cast<uint64_t*>(TR)[w] = 0;
  for (int i = tid; i < cd.n; i += 256) {
    a.AF[cd.row_off + i] = 0;
    a.TM[cd.row_off + i] = 0;
    a.Y[cd.row_off + i] = (double)a.yf[cd.row_off + i];
  }
  __syncthreads();
  for (int j = tid; j < a.m; j += 256) {
    TR[j].ngv[0] = (uint8_t)cd.root_ngv;
    TR[j].used = 1;
    TR[j].leaf = 1;
  }
  if (tid == 0) {
    const uint32_t seed = seeds[chain];
    Rng g{seed ^ 0x9E3779B9u, seed * 2654435761u + 1u, 0x6A09E667u ^ (seed << 7), 0xBB67AE85u + seed};
    for (int i = 0; i < 16; i++) rng_next(g);
    if (a.probit && !a.pr_ext) {
      const double off = (double)cd.offset;
      const uint8_t* YB = a.ybin + cd.row_off;
      double* Y = a.Y + cd.row_off;
      for (int i = 0; i < cd.n; i++) Y[i] = YB[i] == 0 ? -rtnorm_dev(g, 0., off, a.tb) : rtnorm_dev(g, 0., -off, a.tb);
    }
    ChainState* cs = a.state + chain;
    cs->rng[0] = g.s0;
    cs->rng[1] = g.s1;
    cs->rng[2] = g.s2;
    cs->rng[3] = g.s3;
    cs->sigma = (double)cd.sigest;
    cs->it = 0;
    cs->n_birth = cs->n_death = cs->n_cap = 0;
    cs->xseq = 0;
    cs->xerr = 0;
  }
}

// trmean = (float)TM / (float)ndpost (:573-578)
__global__ void k_trmean(const double* TM, float* out, int n, int ndpost) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)TM[i] / (float)ndpost;
}

// ------------------------------------------------------------------ probit latent redraw spread over the device
// The draws of probit_latents (:543-547, rtnorm :282-292) for batches of few chains, where one workgroup per chain
// would leave the other CUs idle. Row i's attempts use slots S_i, S_i + 1, ... of the stream (two outputs a slot);
// with d_i = S_i - i the rejections before row i, d is a monotone path: row i accepts at the first offset d >= d_i
// whose attempt accepts, and d_{i+1} is that offset. Kernels after the sweep kernel (which then skips its own
// redraw), per sweep:
//   k_pr_rows, k_pr_band, k_pr_bandrows: per row, the expected rejections and their variance (geometric, from the
//                row's truncation point); prefix sums (per 256 rows, over those, then within them) give the predicted
//                offset E_i and variance V_i of every row, and its band of offsets [lo_i, lo_i + 64 nw_i), centred on
//                E_i, half-width 3.5 sqrt(V_i) + 32;
//   k_pr_gen:    the stream's first K slots: thread t jumps to slot 64 t with the GF(2) matrices T^(128 2^k) of the
//                generator's state transition (T is linear over GF(2)^128) and steps through 64 slots;
//   k_pr_decide: one wave per row decides every offset of its band (latent_accepts: device math, the host-libm
//                images wherever rounding could matter -- each bit is the reference's decision) into 64-bit masks;
//   k_pr_table:  per block of 64 rows and every entry offset in its first row's band, the offset at which the
//                block's last row accepts, following the decision bits;
//   k_pr_compose, k_pr_chain, k_pr_expand: those tables composed per group of 8 blocks; one wave per chain chains
//                the groups (the entry offset of every group) and leaves the generator state after the last row (a
//                group whose path leaves the tables is walked block by block); one wave per group then gives each
//                of its blocks' entry offsets;
//   k_pr_walk:   one wave per block follows the path from its entry offset through the masks (a ballot resolves
//                the run of rows that accept at the current offset) to each row's accepting slot; an offset outside
//                a row's band, or a slot past K, is decided on the spot (exact; outputs past K are stepped from the
//                last checkpoint);
//   k_pr_value:  each row's latent from its accepting slot (latent_value).
// The bands only decide how much is precomputed: results are the same whatever they are.
constexpr int kJumpLevels = 20;     // thread indices below 2^20: K below 2^26 slots
constexpr int kPrSlots = 64;        // slots per generating thread (and between checkpoints)

// A logical chain of the redraw and its rows in the row pools: one range, or, for a chain whose rows are split over W
// workgroups, W ranges of `share` rows (a multiple of 64: no block of 64 rows straddles two) at a stride of share_pad.
struct PrChain {
  uint64_t row_off;  // pool position of row 0
  int n, share, share_pad, W;
  int dc0;           // device chain (workgroup) of member 0; the members are dc0 .. dc0 + W - 1
  float offset;
  __device__ __forceinline__ uint64_t operator()(int i) const {
    return W == 1 ? row_off + (uint64_t)i
                  : row_off + (uint64_t)(i / share) * (uint64_t)share_pad + (uint64_t)(i % share);
  }
};

struct PrArgs {
  const PrChain* pc;          // [CL] the logical chains
  ChainState* state;
  const double* AF;
  const uint8_t* ybin;
  double* Y;
  LibmTables tb;
  const uint4* jump;          // [kJumpLevels][128]: column j of T^(128 2^k), the state 128 2^k steps after unit state j
  uint2* slots;               // [C][K] the two outputs of every slot
  uint4* ck;                  // [C][K / 64] the generator state before slot 64 t
  float2* ev;                 // [rows] expected rejections of the row's attempts and their variance
  double2* btot;              // [C][nb256] their totals per 256 rows, then the prefix before each 256 rows
  float* E;                   // [rows] predicted offset at each row
  int* lo;                    // [rows] first offset of the row's band
  uint8_t* nw;                // [rows] 64-offset words in the band
  unsigned long long* mask;   // [rows][nwmax] decisions: bit b of word w = the attempt at offset lo + 64 w + b accepts
  uint32_t* acc;              // [rows] accepting slot
  int* tab;                   // [C][nbmax][wmax] per block of 64 rows: exit offset for each entry in its band, or -1
  uint32_t* entry;            // [C][nbmax] the offset at each block's first row
  int* gtab;                  // [C][ngmax][wmax] per group of 8 blocks: the block tables composed
  uint32_t* gentry;           // [C][ngmax] the offset at each group's first row
  unsigned* spot;             // [C] offsets the walk decided on the spot (outside the bands)
  uint32_t K;                 // slots per chain, a multiple of 64
  int nwmax;
  int nbmax, wmax;            // blocks of 64 rows of the longest chain; 64 nwmax
  int nb256;                  // blocks of 256 rows of the longest chain
  int ngmax;                  // groups of 8 blocks of the longest chain
};

// the generator state before slot S: its checkpoint (the last one when S is past K) stepped forward
__device__ __forceinline__ Rng pr_state_at(const uint4* CK, uint32_t K, uint32_t S) {
  const uint32_t q = S < K ? S / kPrSlots : K / kPrSlots - 1;
  const uint4 c = CK[q];
  Rng g{c.x, c.y, c.z, c.w};
  for (uint32_t i = q * kPrSlots; i < S; i++) {
    rng_step(g, 11u, 21u);
    rng_step(g, 11u, 21u);
  }
  return g;
}
// the two outputs of slot k
__device__ __forceinline__ uint2 pr_slot(const uint2* X, const uint4* CK, uint32_t K, uint32_t k) {
  if (k < K) return X[k];
  Rng g = pr_state_at(CK, K, k);
  const uint32_t x1 = starstar(g.s1);
  rng_step(g, 11u, 21u);
  return make_uint2(x1, starstar(g.s1));
}

// expected rejections of a row's attempts and their variance: the attempt rejects with probability rho -- Phi(tau)
// for the normal draw (tau <= 0), else one less the exponential proposal's acceptance sqrt(2 pi) lam
// exp(lam tau - lam^2 / 2) Q(tau) -- so the rejections are geometric: rho / (1 - rho), rho / (1 - rho)^2
__device__ __forceinline__ void pr_rejections(const LatentRow& r, double& e, double& v) {
  double rho = 0.;
  if (r.tau <= 0.)
    rho = 0.5 * erfc(-r.tau * 0.70710678118654752);
  else if (r.tau < 26.)
    rho = 1.0 - 1.2533141373155003 * r.lam * exp(r.lam * r.tau - 0.5 * r.lam * r.lam) *
                    erfc(r.tau * 0.70710678118654752);
  rho = fmin(fmax(rho, 0.), 0.9);
  const double a = 1.0 / (1.0 - rho);
  e = rho * a;
  v = rho * a * a;
}

// exclusive prefix sums of (e, v) over a workgroup of 256 threads; returns the workgroup's totals
__device__ __forceinline__ double2 pr_block_scan(double e, double v, double& pe, double& pv) {
  __shared__ double2 wsum[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double ie = e, iv = v;  // inclusive within the wave
  for (int o = 1; o < 64; o <<= 1) {
    const double te = __shfl_up(ie, o), tv = __shfl_up(iv, o);
    if (lane >= o) {
      ie += te;
      iv += tv;
    }
  }
  if (lane == 63) wsum[wave] = make_double2(ie, iv);
  __syncthreads();
  double be = 0., bv = 0.;
  for (int k = 0; k < wave; k++) {
    be += wsum[k].x;
    bv += wsum[k].y;
  }
  pe = be + ie - e;
  pv = bv + iv - v;
  const double2 tot = make_double2(wsum[0].x + wsum[1].x + wsum[2].x + wsum[3].x,
                                   wsum[0].y + wsum[1].y + wsum[2].y + wsum[3].y);
  return tot;
}

// per row: expected rejections and variance; per 256 rows: their totals
__global__ __launch_bounds__(256) void k_pr_rows(PrArgs A) {
  const int c = blockIdx.y;
  const PrChain& R = A.pc[c];
  if ((int)blockIdx.x * 256 >= R.n) return;
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  double e = 0., v = 0.;
  if (i < R.n) {
    pr_rejections(latent_row(A.AF[R(i)], A.ybin[R(i)], (double)R.offset), e, v);
    A.ev[R(i)] = make_float2((float)e, (float)v);
  }
  double pe, pv;
  const double2 tot = pr_block_scan((double)(float)e, (double)(float)v, pe, pv);
  if (threadIdx.x == 0) A.btot[(size_t)c * A.nb256 + blockIdx.x] = tot;
}

// per chain: exclusive prefix sums of the 256-row totals
__global__ __launch_bounds__(256) void k_pr_band(PrArgs A) {
  const int c = blockIdx.x;
  const int nb = (A.pc[c].n + 255) / 256;
  double2* T = A.btot + (size_t)c * A.nb256;
  double ce = 0., cv = 0.;
  for (int b0 = 0; b0 < nb; b0 += 256) {
    const int b = b0 + (int)threadIdx.x;
    const double2 t = b < nb ? T[b] : make_double2(0., 0.);
    double pe, pv;
    const double2 tot = pr_block_scan(t.x, t.y, pe, pv);
    __syncthreads();
    if (b < nb) T[b] = make_double2(ce + pe, cv + pv);
    ce += tot.x;
    cv += tot.y;
  }
}

// per row: the predicted offset (its 256 rows' offset + the prefix within them) and the band around it
__global__ __launch_bounds__(256) void k_pr_bandrows(PrArgs A) {
  const int c = blockIdx.y;
  const PrChain& R = A.pc[c];
  if ((int)blockIdx.x * 256 >= R.n) return;
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  const float2 x = i < R.n ? A.ev[R(i)] : make_float2(0.f, 0.f);
  double pe, pv;
  (void)pr_block_scan((double)x.x, (double)x.y, pe, pv);
  if (i >= R.n) return;
  const double2 b = A.btot[(size_t)c * A.nb256 + blockIdx.x];
  pe += b.x;
  pv += b.y;
  const double h = 3.5 * sqrt(pv) + 32.;
  const int w = min(A.nwmax, max(3, (int)ceil(h / 32.)));
  A.E[R(i)] = (float)pe;
  A.lo[R(i)] = max(0, (int)floor(pe - 32. * w));
  A.nw[R(i)] = (uint8_t)w;
}

// A quad of lanes per generating thread t (slots 64 t .. 64 t + 63): lane j of the quad applies the jump tables'
// columns of state word j, the quad's XOR combines the four parts; then every lane steps through the 64 slots and
// writes its quarter of them.
__global__ __launch_bounds__(256) void k_pr_gen(PrArgs A) {
  const int c = blockIdx.y, j = threadIdx.x & 3;
  const uint32_t t = blockIdx.x * 64u + (threadIdx.x >> 2);
  const uint32_t nth = A.K / kPrSlots;  // generating threads of a chain
  // levels used by any thread of the chain: the bits of nth - 1; their tables staged in LDS
  const int lev = nth > 1 ? 32 - __builtin_clz(nth - 1) : 0;
  __shared__ uint4 J[kJumpLevels][128];
  for (int k = threadIdx.x; k < lev * 128; k += 256) J[k / 128][k % 128] = A.jump[k];
  __syncthreads();
  if (t >= nth) return;  // a quad's four lanes share t: they leave together, and its shuffles see only live lanes
  const ChainState* cs = A.state + A.pc[c].dc0;
  uint32_t s0 = cs->rng[0], s1 = cs->rng[1], s2 = cs->rng[2], s3 = cs->rng[3];
  for (int k = 0; k < lev; k++) {
    const uint32_t x = j == 0 ? s0 : (j == 1 ? s1 : (j == 2 ? s2 : s3));
    uint32_t o0 = 0, o1 = 0, o2 = 0, o3 = 0;
#pragma unroll
    for (int g = 0; g < 2; g++) {
      uint4 col[16];
#pragma unroll
      for (int q = 0; q < 16; q++) col[q] = J[k][32 * j + 16 * g + q];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const uint32_t msk = 0u - ((x >> (16 * g + q)) & 1u);
        o0 ^= col[q].x & msk;
        o1 ^= col[q].y & msk;
        o2 ^= col[q].z & msk;
        o3 ^= col[q].w & msk;
      }
    }
    // the quad's four parts
    o0 ^= __shfl_xor(o0, 1);
    o1 ^= __shfl_xor(o1, 1);
    o2 ^= __shfl_xor(o2, 1);
    o3 ^= __shfl_xor(o3, 1);
    o0 ^= __shfl_xor(o0, 2);
    o1 ^= __shfl_xor(o1, 2);
    o2 ^= __shfl_xor(o2, 2);
    o3 ^= __shfl_xor(o3, 2);
    const bool take = (t >> k) & 1u;
    s0 = take ? o0 : s0;
    s1 = take ? o1 : s1;
    s2 = take ? o2 : s2;
    s3 = take ? o3 : s3;
  }
  Rng g{s0, s1, s2, s3};
  if (j == 0) A.ck[(size_t)c * nth + t] = make_uint4(g.s0, g.s1, g.s2, g.s3);
  uint2* out = A.slots + (size_t)c * A.K + (size_t)t * kPrSlots;
  for (int k = 0; k < kPrSlots; k++) {
    const uint32_t x1 = starstar(g.s1);
    rng_step(g, 11u, 21u);
    const uint32_t x2 = starstar(g.s1);
    rng_step(g, 11u, 21u);
    if ((k >> 4) == j) out[k] = make_uint2(x1, x2);
  }
}

__global__ __launch_bounds__(256) void k_pr_decide(PrArgs A) {
  const int c = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const PrChain& R = A.pc[c];
  const int i = (int)blockIdx.x * 4 + wave;
  if (i >= R.n) return;
  const LatentRow r = latent_row(A.AF[R(i)], A.ybin[R(i)], (double)R.offset);
  const int lo = A.lo[R(i)], nw = A.nw[R(i)];
  const uint2* X = A.slots + (size_t)c * A.K;
  unsigned long long* M = A.mask + R(i) * (uint64_t)A.nwmax;
  for (int w = 0; w < nw; w++) {
    const uint32_t slot = (uint32_t)(i + lo + 64 * w + lane);
    bool ok = false;
    if (slot < A.K) {
      const uint2 x = X[slot];
      ok = latent_accepts(r, x.x, x.y, A.tb);
    }
    const unsigned long long b = __ballot(ok);
    if (lane == 0) M[w] = b;
  }
}

// A lane's view of its row in a window of 64 rows: the row's predicted offset and band (pr_meta), then three mask
// words around the offset predicted for the row from a known offset rb at the row whose predicted offset is eb
// (pr_window), and the offsets they decide, [vlo, vhi).
struct PrMeta {
  float e;
  int lo, nw;
};
struct PrWin {
  unsigned long long w0, w1, w2;
  int vlo, vhi;
};
__device__ __forceinline__ PrMeta pr_meta(const PrArgs& A, const PrChain& R, int n, int i0, int lane) {
  PrMeta M{0.f, 0, 3};
  const int i = i0 + lane;
  if (i < n) {
    M.e = A.E[R(i)];
    M.lo = A.lo[R(i)];
    M.nw = A.nw[R(i)];
  }
  return M;
}
__device__ __forceinline__ PrWin pr_window(const PrArgs& A, const PrMeta& M, const PrChain& R, int n, int i0,
                                           uint32_t rb, float eb, int lane) {
  PrWin W{0, 0, 0, 0, 0};
  const int i = i0 + lane;
  if (i < n) {
    const float pred = (float)rb + (M.e - eb);
    const int q = min(max((int)floorf((pred - (float)M.lo) * (1.f / 64.f)) - 1, 0), M.nw - 3);
    const unsigned long long* P = A.mask + R(i) * (uint64_t)A.nwmax + q;
    W.w0 = P[0];
    W.w1 = P[1];
    W.w2 = P[2];
    W.vlo = M.lo + 64 * q;
    W.vhi = min(W.vlo + 192, (int)A.K - i);
  }
  return W;
}

// row i's first accepting offset at or above `from`, decided here: 64 offsets at a time across the wave. The bound
// only guarantees an exit (an attempt rejects with probability at most 1/2: 2^22 rejections in a row do not occur).
__device__ uint32_t pr_spot(const PrArgs& A, const uint2* X, const uint4* CK, const PrChain& R, double off, int i,
                            uint32_t from, int lane) {
  const LatentRow r = latent_row(A.AF[R(i)], A.ybin[R(i)], off);
  uint32_t o = from;
  for (int k = 0; k < (1 << 16); k++, o += 64) {
    const uint2 x = pr_slot(X, CK, A.K, (uint32_t)i + o + (uint32_t)lane);
    const unsigned long long b = __ballot(latent_accepts(r, x.x, x.y, A.tb));
    if (b) return o + (uint32_t)__builtin_ctzll(b);
  }
  return o;
}

// Rows [i0, i0 + reff) from offset r, with each lane's words around the path: the run of rows that accept at the
// current offset is resolved by one ballot, then the first row that does not moves the offset to its next accepting
// one (from its words, or decided on the spot). Writes each row's accepting slot if `write`; returns the offset at
// which the last row accepts (the next row's starting offset).
__device__ __forceinline__ uint32_t pr_walk_window(const PrArgs& A, const PrWin& cur, const uint2* X, const uint4* CK,
                                                   const PrChain& R, double off, int i0, int reff, uint32_t r,
                                                   unsigned& spot, bool write, int lane) {
  const unsigned long long live = reff >= 64 ? ~0ull : ((1ull << reff) - 1ull);
  int done = 0;
  uint32_t accv = 0;
  while (done < reff) {
    // every lane: its row's first accepting offset at or above r from its words, or -1 - (the first offset not known
    // to reject) when they do not reach one
    const int rel = (int)r - cur.vlo, lim = cur.vhi - cur.vlo;
    int nx = -1 - (int)r;
    if (rel >= 0 && rel < lim) {
      const unsigned long long m0 = rel < 64 ? (cur.w0 & (~0ull << rel)) : 0ull;
      const unsigned long long m1 = rel < 64 ? cur.w1 : (rel < 128 ? (cur.w1 & (~0ull << (rel - 64))) : 0ull);
      const unsigned long long m2 = rel < 128 ? cur.w2 : (cur.w2 & (~0ull << (rel - 128)));
      const int pos = m0 ? (int)__builtin_ctzll(m0)
                         : (m1 ? 64 + (int)__builtin_ctzll(m1) : (m2 ? 128 + (int)__builtin_ctzll(m2) : 192));
      nx = pos < lim ? cur.vlo + pos : -1 - cur.vhi;
    }
    const bool ok = nx == (int)r;
    const unsigned long long notok = live & (~0ull << done) & ~(unsigned long long)__ballot(ok);
    const int f = notok ? (int)__builtin_ctzll(notok) : reff;
    accv = (lane >= done && lane < f) ? r : accv;
    done = f;
    if (done >= reff) break;
    const int nf = __builtin_amdgcn_readlane(nx, f);
    if (nf >= 0) {
      r = (uint32_t)nf;
    } else {
      r = pr_spot(A, X, CK, R, off, i0 + f, (uint32_t)(-1 - nf), lane);
      spot++;
    }
    accv = lane == f ? r : accv;
    done = f + 1;
  }
  if (write && lane < reff) A.acc[R(i0 + lane)] = (uint32_t)(i0 + lane) + accv;
  return r;
}

// Per block of 64 rows and every entry offset in the band of its first row: the offset at which the block's last
// row accepts, following the decision bits (-1 where the path leaves a row's band). Bands staged in LDS.
__global__ __launch_bounds__(256) void k_pr_table(PrArgs A) {
  const int c = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  const PrChain& R = A.pc[c];
  const int n = R.n, r0 = 64 * b;
  if (r0 >= n) return;
  const int nr = min(64, n - r0), nwm = A.nwmax;
  extern __shared__ unsigned long long pr_sm[];  // [64][nwmax] band words, then lo[64], nw[64]
  int* slo = reinterpret_cast<int*>(pr_sm + 64 * nwm);
  int* snw = slo + 64;
  // the block's rows are consecutive in the pools (a block never straddles two members of a split chain)
  for (int k = tid; k < nr * nwm; k += 256) pr_sm[k] = A.mask[R(r0) * (uint64_t)nwm + k];
  for (int k = tid; k < nr; k += 256) {
    slo[k] = A.lo[R(r0 + k)];
    snw[k] = A.nw[R(r0 + k)];
  }
  __syncthreads();
  const int L = slo[0], Wd = 64 * snw[0];
  int* T = A.tab + ((size_t)c * A.nbmax + b) * (size_t)A.wmax;
  // kPrTabE entries per thread walked together: independent chains of dependent LDS reads, issued side by side
  constexpr int kPrTabE = 4;
  for (int u0 = tid; u0 < Wd; u0 += 256 * kPrTabE) {
    int x[kPrTabE];
#pragma unroll
    for (int e = 0; e < kPrTabE; e++) x[e] = u0 + 256 * e < Wd ? L + u0 + 256 * e : -1;
    for (int l = 0; l < nr; l++) {
      const int lo = slo[l], nw = snw[l];
#pragma unroll
      for (int e = 0; e < kPrTabE; e++) {
        if (x[e] < 0) continue;
        const int rel = x[e] - lo;
        if (rel < 0 || rel >= 64 * nw) {
          x[e] = -1;
          continue;
        }
        int wi = rel >> 6;
        unsigned long long m = pr_sm[l * nwm + wi] & (~0ull << (rel & 63));
        while (!m && ++wi < nw) m = pr_sm[l * nwm + wi];
        x[e] = m ? lo + 64 * wi + (int)__builtin_ctzll(m) : -1;
      }
    }
#pragma unroll
    for (int e = 0; e < kPrTabE; e++)
      if (u0 + 256 * e < Wd) T[u0 + 256 * e] = x[e];
  }
}

// Blocks [b0, b1) from entry offset x through the block tables, one dependent lookup a block (every block's band
// start and width fetched first, lane j holding block b0 + j's); a block whose entry is outside its table, or whose
// path leaves the bands, is walked row by row. Stores each block's entry offset if `entry`; returns the exit.
__device__ __forceinline__ uint32_t pr_blocks_walk(const PrArgs& A, int c, int n, const PrChain& R, double off,
                                                   const uint2* X, const uint4* CK, int b0, int b1, uint32_t x,
                                                   uint32_t* entry, unsigned& spot, int lane) {
  const int* TB = A.tab + (size_t)c * A.nbmax * (size_t)A.wmax;
  const int bl = b0 + lane;
  const int mL = lane < b1 - b0 ? A.lo[R(64 * bl)] : 0, mW = lane < b1 - b0 ? 64 * (int)A.nw[R(64 * bl)] : 0;
  for (int b = b0; b < b1; b++) {
    if (entry && lane == 0) entry[b] = x;
    const int t = (int)x - __builtin_amdgcn_readlane(mL, b - b0);
    const int v = (t >= 0 && t < __builtin_amdgcn_readlane(mW, b - b0)) ? TB[(size_t)b * A.wmax + t] : -1;
    if (v >= 0) {
      x = (uint32_t)v;
    } else {
      const int i0 = 64 * b;
      const PrWin cur = pr_window(A, pr_meta(A, R, n, i0, lane), R, n, i0, x, A.E[R(i0)], lane);
      x = pr_walk_window(A, cur, X, CK, R, off, i0, min(64, n - i0), x, spot, false, lane);
    }
  }
  return x;
}

// Per group of kPrGroup blocks and every entry offset in its first block's band: the exit offset of the group, the
// block tables composed (-1 where a path leaves them).
constexpr int kPrGroup = 8;
__global__ __launch_bounds__(256) void k_pr_compose(PrArgs A) {
  const int c = blockIdx.y, g = blockIdx.x, tid = threadIdx.x;
  const PrChain& R = A.pc[c];
  const int n = R.n, nb = (n + 63) / 64, b0 = kPrGroup * g;
  if (b0 >= nb) return;
  const int b1 = min(nb, b0 + kPrGroup);
  __shared__ int sL[kPrGroup], sW[kPrGroup];
  if (tid < b1 - b0) {
    sL[tid] = A.lo[R(64 * (b0 + tid))];
    sW[tid] = 64 * (int)A.nw[R(64 * (b0 + tid))];
  }
  __syncthreads();
  const int* TB = A.tab + (size_t)c * A.nbmax * (size_t)A.wmax;
  int* G = A.gtab + ((size_t)c * A.ngmax + g) * (size_t)A.wmax;
  for (int u = tid; u < sW[0]; u += 256) {
    int x = sL[0] + u;
    for (int b = b0; b < b1 && x >= 0; b++) {
      const int t = x - sL[b - b0];
      x = (t >= 0 && t < sW[b - b0]) ? TB[(size_t)b * A.wmax + t] : -1;
    }
    G[u] = x;
  }
}

// One wave per chain: the entry offset of every group of blocks, group after group through the group tables. Groups
// go kPrAhead at a time: at their start the table entries around each one's predicted entry (fetched during the
// previous kPrAhead) go to LDS and those of the next kPrAhead are fetched; the dependent steps then read LDS only
// (the groups' first-row bands staged in LDS kPrChunk groups at a time). An entry outside the fetched window is
// looked up in HBM; a group whose path leaves the tables is walked block by block. Leaves the generator state after
// the last row.
constexpr int kPrAhead = 4;
constexpr int kPrWin = 256;   // table entries fetched around a group's predicted entry (4 a lane)
constexpr int kPrChunk = 256;
__global__ __launch_bounds__(64) void k_pr_chain(PrArgs A) {
  const int c = blockIdx.x, lane = threadIdx.x;
  const PrChain& R = A.pc[c];
  const int n = R.n, nb = (n + 63) / 64, ng = (nb + kPrGroup - 1) / kPrGroup;
  const double off = (double)R.offset;
  const uint2* X = A.slots + (size_t)c * A.K;
  const uint4* CK = A.ck + (size_t)c * (A.K / kPrSlots);
  const int* GT = A.gtab + (size_t)c * A.ngmax * (size_t)A.wmax;
  uint32_t* gentry = A.gentry + (size_t)c * A.ngmax;
  // first row of groups [q0, q0 + kPrChunk + 2 kPrAhead): predicted offset, band start, band width
  __shared__ float sE[kPrChunk + 2 * kPrAhead];
  __shared__ int sL[kPrChunk + 2 * kPrAhead], sW[kPrChunk + 2 * kPrAhead];
  // the current groups' table windows: kPrWin entries from swu[j]
  __shared__ __align__(16) int win[kPrAhead][kPrWin];
  __shared__ int swu[kPrAhead];
  int q0 = 0;
  auto stage = [&](int q) {
    for (int k = lane; k < kPrChunk + 2 * kPrAhead; k += 64) {
      const int gi = q + k;
      const uint64_t r = gi < ng ? R(64 * kPrGroup * gi) : 0;
      sE[k] = gi < ng ? A.E[r] : 0.f;
      sL[k] = gi < ng ? A.lo[r] : 0;
      sW[k] = gi < ng ? 64 * (int)A.nw[r] : 0;
    }
    wave_sync();
  };
  // lane k: entries u0 + 4 k .. u0 + 4 k + 3 of group gi's table
  auto fetch = [&](int gi, int u0, int W, int4& t) {
    const int u = u0 + 4 * lane;
    const int* T = GT + (size_t)gi * A.wmax;
    t.x = (gi < ng && u >= 0 && u < W) ? T[u] : -2;
    t.y = (gi < ng && u + 1 >= 0 && u + 1 < W) ? T[u + 1] : -2;
    t.z = (gi < ng && u + 2 >= 0 && u + 2 < W) ? T[u + 2] : -2;
    t.w = (gi < ng && u + 3 >= 0 && u + 3 < W) ? T[u + 3] : -2;
  };
  stage(0);
  int4 tq[kPrAhead];
  int u0v[kPrAhead];
#pragma unroll
  for (int j = 0; j < kPrAhead; j++) {
    u0v[j] = (int)rintf(sE[j] - sE[0]) - sL[j] - kPrWin / 2;
    fetch(j, u0v[j], sW[j], tq[j]);
  }
  uint32_t x = 0;
  unsigned spot = 0;
  for (int g0 = 0; g0 < ng; g0 += kPrAhead) {
    if (g0 >= q0 + kPrChunk) {
      q0 = g0;
      stage(q0);
    }
    // these groups' windows to LDS (fetched kPrAhead groups ago), the next ones fetched from here
#pragma unroll
    for (int j = 0; j < kPrAhead; j++) {
      *reinterpret_cast<int4*>(&win[j][4 * lane]) = tq[j];
      if (lane == j) swu[j] = u0v[j];
    }
    wave_sync();
    const float eg0 = sE[g0 - q0];
#pragma unroll
    for (int j = 0; j < kPrAhead; j++) {
      const int gn = g0 + kPrAhead + j, k = gn - q0;
      if (gn < ng) {
        u0v[j] = (int)x + (int)rintf(sE[k] - eg0) - sL[k] - kPrWin / 2;
        fetch(gn, u0v[j], sW[k], tq[j]);
      }
    }
    // lane j holds group g0 + j's band start, band width and window start
    const int gL = lane < kPrAhead ? sL[g0 - q0 + lane] : 0, gW = lane < kPrAhead ? sW[g0 - q0 + lane] : 0;
    const int gU = lane < kPrAhead ? swu[lane] : 0;
#pragma unroll 1
    for (int j = 0; j < kPrAhead; j++) {
      const int gi = g0 + j;
      if (gi < ng) {
        if (lane == 0) gentry[gi] = x;
        const int u = (int)x - __builtin_amdgcn_readlane(gL, j);
        int v = -1;
        if (u >= 0 && u < __builtin_amdgcn_readlane(gW, j)) {
          const int kk = u - __builtin_amdgcn_readlane(gU, j);
          v = (kk >= 0 && kk < kPrWin) ? win[j][kk] : GT[(size_t)gi * A.wmax + u];
        }
        if (v >= 0)
          x = (uint32_t)v;
        else
          x = pr_blocks_walk(A, c, n, R, off, X, CK, kPrGroup * gi, min(nb, kPrGroup * gi + kPrGroup), x, nullptr,
                             spot, lane);
      }
    }
    wave_sync();
  }
  // every member of a split chain continues from the same generator state
  if (lane < R.W) {
    const Rng g = pr_state_at(CK, A.K, (uint32_t)n + x);
    ChainState* cs = A.state + R.dc0 + lane;
    cs->rng[0] = g.s0;
    cs->rng[1] = g.s1;
    cs->rng[2] = g.s2;
    cs->rng[3] = g.s3;
  }
}

// One wave per group: the entry offset of each of its blocks, from the group's.
__global__ __launch_bounds__(64) void k_pr_expand(PrArgs A) {
  const int c = blockIdx.y, g = blockIdx.x, lane = threadIdx.x;
  const PrChain& R = A.pc[c];
  const int n = R.n, nb = (n + 63) / 64, b0 = kPrGroup * g;
  if (b0 >= nb) return;
  unsigned spot = 0;
  (void)pr_blocks_walk(A, c, n, R, (double)R.offset, A.slots + (size_t)c * A.K,
                       A.ck + (size_t)c * (A.K / kPrSlots), b0, min(nb, b0 + kPrGroup),
                       A.gentry[(size_t)c * A.ngmax + g], A.entry + (size_t)c * A.nbmax, spot, lane);
}

// One wave per block of 64 rows: each row's accepting slot, from the block's entry offset.
__global__ __launch_bounds__(64) void k_pr_walk(PrArgs A) {
  const int c = blockIdx.y, b = blockIdx.x, lane = threadIdx.x;
  const PrChain& R = A.pc[c];
  const int n = R.n, i0 = 64 * b;
  if (i0 >= n) return;
  const uint32_t r = A.entry[(size_t)c * A.nbmax + b];
  const PrWin cur = pr_window(A, pr_meta(A, R, n, i0, lane), R, n, i0, r, A.E[R(i0)], lane);
  unsigned spot = 0;
  (void)pr_walk_window(A, cur, A.slots + (size_t)c * A.K, A.ck + (size_t)c * (A.K / kPrSlots), R,
                       (double)R.offset, i0, min(64, n - i0), r, spot, true, lane);
  if (lane == 0 && spot) atomicAdd(A.spot + c, spot);
}

__global__ __launch_bounds__(256) void k_pr_value(PrArgs A) {
  const int c = blockIdx.y;
  const PrChain& R = A.pc[c];
  const int i = (int)blockIdx.x * 256 + (int)threadIdx.x;
  if (i >= R.n) return;
  const uint2 x = pr_slot(A.slots + (size_t)c * A.K, A.ck + (size_t)c * (A.K / kPrSlots), A.K, A.acc[R(i)]);
  const LatentRow r = latent_row(A.AF[R(i)], A.ybin[R(i)], (double)R.offset);
  A.Y[R(i)] = latent_value(r, x.x, x.y, A.tb);
}

// ------------------------------------------------------------------ launch variants
using SweepKernel = void (*)(KArgs);
constexpr int kRowsRes = 6;    // 1536-row tiles (the running fit resident in LDS)
constexpr int kRowsStream = 4; // 1024-row tiles (the running fit streamed, or resident in LDS)
static_assert(sizeof(ProbitRing) <= Geo<kRowsStream>::kScratchBytes && sizeof(ProbitRing) <= Geo<kRowsRes>::kScratchBytes,
              "the probit ring reuses the tile scratch");
// tile kinds: 0 = 1024-row tiles, running fit streamed; 1 = 1024-row tiles, fit in LDS; 2 = 1536-row tiles, fit in LDS
constexpr int kKinds = 3;
constexpr bool kind_resident(int k) { return k != 0; }
constexpr int kind_rows(int k) { return k == 2 ? kRowsRes : kRowsStream; }
inline int kind_scratch(int k) { return k == 2 ? Geo<kRowsRes>::kScratchBytes : Geo<kRowsStream>::kScratchBytes; }
constexpr int kCtlBytes = sizeof(Ctl);
constexpr size_t kLdsBudgetPerWG = 80 * 1024;
inline size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

struct Variant {
  SweepKernel fn;
  int scratch;
  int kind;
  bool resident;
  bool dense;
};

template <int R, bool RES, bool PB, bool D>
inline Variant variant_of(int kind, bool fx) {
  return fx ? Variant{k_sweeps<R, RES, true, PB, D>, Geo<R>::kScratchBytes, kind, RES, D}
            : Variant{k_sweeps<R, RES, false, PB, D>, Geo<R>::kScratchBytes, kind, RES, D};
}
template <bool PB, bool D>
inline Variant pick_variant_pd(int kind, bool fx) {
  if (kind == 2) return variant_of<kRowsRes, true, PB, D>(kind, fx);
  if (kind == 1) return variant_of<kRowsStream, true, PB, D>(kind, fx);
  return variant_of<kRowsStream, false, PB, D>(kind, fx);
}
inline Variant pick_variant(int kind, bool fx, bool probit, bool dense) {
  if (probit) return dense ? pick_variant_pd<true, true>(kind, fx) : pick_variant_pd<true, false>(kind, fx);
  return dense ? pick_variant_pd<false, true>(kind, fx) : pick_variant_pd<false, false>(kind, fx);
}

constexpr int kRowsStreamL = kRowsStream;  // the learned pipeline streams the running fit

// ==================================================================== learned birth proposal: lock-step pipeline
// Per tree update j, over all chains of the batch: stage 1 (apply tree j-1's change, leaf sums of tree j, move type
// and node, the node's features), the network's forward pass on the matrix cores (fbh_mlp.hip), stage 2 (the rule
// from the mixture (1 - alpha) uniform + alpha network, split sums, Metropolis-Hastings with the proposal correction,
// leaf values). The proposal correction (no reference counterpart; it plugs into the ratios of
// src/bart_kernel.cpp:451 and :462): a birth with rule (v, c) multiplies pr by q_unif / q_mix, a death of a node with
// rule (v, c) by q_mix / q_unif, where q_unif = (1/p) / (U - L + 1) is BART's rule draw and q_mix = (1 - alpha) q_unif +
// alpha q_net(v, c | features of the node). The features come from the node's partial residuals (excluding tree j),
// computed from an int64 fixed-point running fit so that they are bitwise the same in the state before a birth and in
// the state after it (where the reverse death evaluates them; sigma is fixed within a sweep, and the node's rows and
// cut ranges are the same in both states). Per variable v with cut range [L, U] at the node, 16 cut positions
// log-spaced from both ends of the range (L + 2^k - 1 and U - 2^k + 1, k = 0..7) split the node's rows into 17
// segments, each cut further at every 16th rank (bin of rank r: the positions below r plus r / 16, at most 31, which
// keeps the lanes of a wave on different histogram words); the node's histograms of w and w r over those bins give,
// at each position c, the left child's sums
// (W_c, S_c) and the split's standardized log-likelihood gain
//   g = z^2 / 2 - log(f (1 - f)) / 2,  f = W_c / W,  z = (S_c - f S) / sqrt(W f (1 - f)) / sigma
// (W, S the node's totals): the large-sum form of lh(left) + lh(right) - lh(node) (src/bart_kernel.cpp:278) up to
// terms equal for every split. Positions outside [L, U] give 0, the last feature is the node's depth / 8. The
// likelihood gain concentrates the split conditional near the range ends, where log-spaced positions resolve it; a
// plain 16-bin residual histogram over ranks (rank >> 4) leaves the network at the level of the class prior.
constexpr int kFeatPos = 16;                           // cut positions per variable
constexpr int kSegs = 32;                              // histogram bins per variable
constexpr int kFeat = FBH_PMAX * kFeatPos + 1;         // 513: gains + depth
constexpr int kCuts = 255;                             // classes (v, c) -> v * 255 + c
constexpr int kClasses = FBH_PMAX * kCuts;             // 8160
constexpr double kFxFit = 4294967296.0;                // running fit at 2^-32
constexpr double kFxFeat = 1048576.0;                  // feature sums at 2^-20
constexpr uint16_t kNoRange = 0x00FF;                  // L = 255 > U = 0
constexpr int kHistCopies = 2;                         // copies of the node histograms (a power of two)

constexpr double kFxW = 4096.0;                        // weight histogram at 2^-12 (uint32: a node's total < 2^31)

// position k of variable v's 16 at a node with cut range [L, U] (U < L: none); -1 when outside the range
__device__ __forceinline__ int feat_pos(int L, int U, int k) {
  if (U < L) return -1;
  const int c = k < 8 ? L + (1 << k) - 1 : U - (1 << (15 - k)) + 1;
  return (c >= L && c <= U) ? c : -1;
}
// number of k with 2^k <= x, k = 0..7 (x >= 0)
__device__ __forceinline__ int pow2_upto(int x) { return x > 0 ? min(8, 32 - __clz(x)) : 0; }
// bin of cutpoint rank r: r / 16 plus the positions below r. The positions below r from the left end are the k with
// L + 2^k - 1 < r and 2^k <= U - L + 1; from the right end the k with U - 2^k + 1 < r and 2^k <= U - L + 1.
__device__ __forceinline__ int feat_bin(int L, int U, int r) {
  if (U < L) return r >> 4;
  const int K = U - L + 1;
  const int left = pow2_upto(min(r - L, K));
  const int right = pow2_upto(K) - pow2_upto(U + 1 - r);
  return (r >> 4) + left + max(right, 0);
}

struct LpState {
  uint32_t rng[4];
  double sigma;
  int it, mode, nx, n_birth, n_death, n_cap;
};

struct LpArgs {
  KArgs a;
  Ctl* img;                 // [C] control images
  LpState* st;              // [C]
  long long* AFX;           // fixed-point running fit (row pools)
  float* X;                 // [C][kFeat] features of each chain's proposal node
  uint16_t* rg;             // [C][32] cutpoint ranges at the node, L | U << 8
  const float* Z;           // network logits [>= C][ldz]
  int ldz;
  float alpha;
  float* px;                // training pairs: features [pcap][kFeat], ranges [pcap][32], class [pcap]
  uint16_t* prg;
  int* ptg;
  unsigned long long* pcount;
  long pcap;
  float* pent;              // [pcap][2]: entropy of the pair's split conditional, log of its allowed class count
  float* pq;                // [pcap][kClasses]: the pair's split conditional (target kind 1)
  unsigned long long tseed; // label and selection draws of the conditional target (apart from the chains' generators)
  float keep;               // target kind 1: share of the proposed births taken as pairs
  int target;               // pairs: 0 = accepted births with BART's rule, 1 = proposed births, conditional target
  int j;
};

__device__ __forceinline__ void img_copy(void* dst, const void* src, int tid) {
  const uint4* s = reinterpret_cast<const uint4*>(src);
  uint4* d = reinterpret_cast<uint4*>(dst);
  for (int w = tid; w < (int)(sizeof(Ctl) / 16); w += kBlock) d[w] = s[w];
}
static_assert(sizeof(Ctl) % 16 == 0, "control image copied in 16-byte words");

__device__ __forceinline__ long long fit_fx(double th) { return (long long)rint(th * kFxFit); }

__device__ __forceinline__ RowPtrs lp_rows(const KArgs& a, const ChainDesc& cd) {
  RowPtrs P;
  P.AF = a.AF + cd.row_off;
  P.AFl = nullptr;
  P.TM = a.TM + cd.row_off;
  P.Y = a.Y + cd.row_off;
  P.yf = a.yf + cd.row_off;
  P.W = a.W + cd.row_off;
  P.Q = a.q + cd.q_off;
  P.LEAF = a.leaf + cd.leaf_off;
  P.n = cd.n;
  P.probit = 0;
  P.ns = cd.ns;
  return P;
}

// the fixed-point fit follows the change of tree jp (c.tho -> c.thn, leaf ids jp before the data pass rewrites
// them). Each data wave takes the row quads of its own summation lanes, as the data pass does (rows 16 s + 4 dw ..
// 16 s + 4 dw + 3: one dword of each byte column, 16-byte aligned and padded), the loads of a quad issued together.
__device__ void lp_afx_apply(const Ctl& c, const RowPtrs& P, long long* __restrict__ AFX, int jp, int dw, int lane) {
  const int n = P.n;
  const int pmode = c.pmode, pnx = c.pnx, pc = c.pc, pL = c.pL, pR = c.pR, pl = c.pl, pr = c.pr;
  const uint8_t* lp = P.LEAF + (size_t)jp * P.ns;
  const uint8_t* qp = P.Q + (size_t)(pmode == 1 ? c.pv : 0) * P.ns;
  for (int i4 = 16 * lane + kLanesPerWave * dw; i4 < n; i4 += 16 * 64) {
    const uint32_t lw = *reinterpret_cast<const uint32_t*>(lp + i4);
    const uint32_t qw = *reinterpret_cast<const uint32_t*>(qp + i4);
    long long x[4];
#pragma unroll
    for (int u = 0; u < 4; u++) x[u] = i4 + u < n ? AFX[i4 + u] : 0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if (i4 + u >= n) break;
      const int lf = (int)((lw >> (8 * u)) & 255u);
      int lf2 = lf;
      if (pmode == 1 && lf == pnx)
        lf2 = (int)((qw >> (8 * u)) & 255u) <= pc ? pL : pR;
      else if (pmode == 2 && (lf == pl || lf == pr))
        lf2 = pnx;
      AFX[i4 + u] = x[u] + (fit_fx(c.thn[lf2]) - fit_fx(c.tho[lf]));
    }
  }
}

// leaf tables of tree buffer 0 (rank / slot of the used leaves, values) and the move type and node (:413-431, :455)
__device__ void lp_choose(Ctl& c, Rng& g, int lane, int* mode_out, int* nx_out) {
  const TreeDev& t = c.tree[0];
  const uint64_t used = t.used, leafm = t.leaf;
  const int s = lane;
  const bool isleaf = BITM(leafm, s), isused = BITM(used, s);
  const bool good = isused && isleaf && t.ngv[s] > 0;
  const bool nog = isused && !isleaf && BITM(leafm, t.l[s] & 63) && BITM(leafm, t.r[s] & 63);
  const uint64_t goodm = __ballot(good), nogm = __ballot(nog);
  const uint64_t lm = used & leafm;
  c.thr[0][s].th = (double)t.th[s];
  if (BITM(lm, s)) {
    const int rk = __popcll(lm & ((1ull << s) - 1ull));
    c.thr[0][s].rank = (uint32_t)rk;
    c.slot_of[0][rk] = (uint8_t)s;
  } else {
    c.thr[0][s].rank = 0;
  }
  const int ngood = __popcll(goodm), nnogs = __popcll(nogm);
  const double PBx = ngood == 0 ? 0 : (used == 1ull ? 1 : 0.5);
  // the same draws as the plain proposal (:431, :433, :455): a death is drawn whenever a birth is not
  int mode, nx;
  if (unif_d(g) < PBx) {
    mode = 1;
    nx = select_kth(goodm, (int)(unif_d(g) * ngood));
  } else {
    mode = 2;
    nx = select_kth(nogm, (int)(unif_d(g) * nnogs));
  }
  if (lane == 0) {
    c.nleaf[0] = __popcll(lm);
    c.ngood[0] = ngood;
    c.nnogs[0] = nnogs;
    c.cmode[0] = 0;
  }
  *mode_out = mode;
  *nx_out = nx;
}

template <bool FX>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(4))) void k_lp_stage1(LpArgs A) {
  __shared__ Ctl c;
  // the node's histograms of w (uint32 at 2^-12) and w r (int64 at 2^-20) over the row bins of each variable, in
  // kHistCopies copies, lane l adding into copy l % kHistCopies (fewer lanes serialised on one address); the copies
  // are added after
  __shared__ unsigned H0[kHistCopies][FBH_PMAX][kSegs];               // w at 2^-12
  __shared__ unsigned long long H1[kHistCopies][FBH_PMAX][kSegs];     // w r at 2^-20
  __shared__ uint8_t seg[FBH_PMAX][256];  // bin of each cutpoint rank: the positions below it plus rank / 16
  __shared__ uint16_t rgs[FBH_PMAX];      // cut ranges at the node, L | U << 8
  unsigned char* const dyn = fbh_dyn;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, dw = wave - 1;
  const int chain = blockIdx.x, j = A.j;
  const KArgs& a = A.a;
  img_copy(&c, A.img + chain, tid);
  for (int k = tid; k < kHistCopies * FBH_PMAX * kSegs; k += kBlock) {
    (&H0[0][0][0])[k] = 0;
    (&H1[0][0][0])[k] = 0;
  }
  __syncthreads();
  const RowPtrs P = lp_rows(a, c.cd);
  const int n = c.cd.n, p = c.cd.p;
  TreeDev* TR = a.trees + c.cd.tree_off;
  LpState* S = A.st + chain;
  long long* AFX = A.AFX + c.cd.row_off;
  if (wave == 0) {
    if (j == 0 && lane == 0) {
      const double sg = S->sigma;
      c.inv_s2 = 1.0 / (sg * sg);
      c.pmode = 0;
    }
    copy_tree(&c.tree[0], &TR[j], lane);
    wave_sync();
    Rng g{ufl(S->rng[0]), ufl(S->rng[1]), ufl(S->rng[2]), ufl(S->rng[3])};
    int mode, nx;
    lp_choose(c, g, lane, &mode, &nx);
    if (lane == 0) {
      S->rng[0] = g.s0;
      S->rng[1] = g.s1;
      S->rng[2] = g.s2;
      S->rng[3] = g.s3;
      S->mode = mode;
      S->nx = nx;
      c.cnx[0] = nx;
      c.cv[0] = mode;  // stage 1 keeps the move type here until stage 2 sets the rule
    }
    if (lane < FBH_PMAX) {
      uint16_t rv = kNoRange;
      if (mode && lane < p) {
        int L, U;
        node_range(c.tree[0], nx, lane, c.cd.nc[lane], &L, &U);
        if (U >= L) rv = (uint16_t)(L | (U << 8));
      }
      rgs[lane] = rv;
    }
  }
  __syncthreads();
  if (wave > 0) {
    if (j > 0) lp_afx_apply(c, P, AFX, j - 1, dw, lane);
    TileIn<kRowsStreamL> T;
    data_pass<kRowsStreamL, false, FX, true, false>(c, dyn, P, dw, lane, j - 1, j, 0, j > 0, true, false, T, false);
  }
  __syncthreads();
  if (wave == 0) ctl_combine<FX>(c, 0, false, 0, lane);
  const int mode = c.cv[0], nx = c.cnx[0];
  if (mode) {  // the same value in every thread
    for (int e = tid; e < p * 256; e += kBlock) {
      const int v = e >> 8, r = e & 255;
      const int L = rgs[v] & 255, U = rgs[v] >> 8;
      seg[v][r] = (uint8_t)feat_bin(L, U, r);
    }
    __syncthreads();
    // the node's rows: the leaf nx (birth) or the two children of nx (death)
    const TreeDev& t = c.tree[0];
    const int l = mode == 2 ? t.l[nx] : nx, r = mode == 2 ? t.r[nx] : nx;
    const uint8_t* lc = P.LEAF + (size_t)j * P.ns;
    auto& H0c = H0[lane & (kHistCopies - 1)];
    auto& H1c = H1[lane & (kHistCopies - 1)];
    // four consecutive rows per thread and iteration: the byte columns (16-byte aligned, padded to a multiple of 16
    // rows) as one dword each, and the cutpoint ranks of four variables loaded together before their adds, so that
    // an iteration waits for memory a few times instead of once per variable and row
    for (int i4 = 4 * tid; i4 < n; i4 += 4 * kBlock) {
      const uint32_t lw = *reinterpret_cast<const uint32_t*>(lc + i4);
      bool in[4];
      unsigned wf[4];
      unsigned long long wr[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int i = i4 + u;
        const int lf = (int)((lw >> (8 * u)) & 255u);
        in[u] = i < n && (lf == l || lf == r);
        wf[u] = wr[u] = 0;
        if (in[u]) {
          const double res = (double)P.Y[i] - (double)(AFX[i] - fit_fx(c.thr[0][lf].th)) * (1.0 / kFxFit);
          const double w = (double)P.W[i];
          wf[u] = (unsigned)rint(w * kFxW);
          wr[u] = (unsigned long long)(long long)rint(w * res * kFxFeat);
        }
      }
      if (!(in[0] || in[1] || in[2] || in[3])) continue;
      for (int v0 = 0; v0 < p; v0 += 4) {
        uint32_t qw[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
          qw[k] = v0 + k < p ? *reinterpret_cast<const uint32_t*>(P.Q + (size_t)(v0 + k) * P.ns + i4) : 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (v0 + k >= p) break;
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (in[u]) {
              const int sg = seg[v0 + k][(qw[k] >> (8 * u)) & 255u];
              atomicAdd(&H0c[v0 + k][sg], wf[u]);
              atomicAdd(&H1c[v0 + k][sg], wr[u]);
            }
        }
      }
    }
  }
  __syncthreads();
  if (mode)
    for (int k = tid; k < FBH_PMAX * kSegs; k += kBlock) {
      unsigned t0 = (&H0[0][0][0])[k];
      unsigned long long t1 = (&H1[0][0][0])[k];
      for (int h = 1; h < kHistCopies; h++) {
        t0 += (&H0[h][0][0])[k];
        t1 += (&H1[h][0][0])[k];
      }
      (&H0[0][0][0])[k] = t0;
      (&H1[0][0][0])[k] = t1;
    }
  __syncthreads();
  float* X = A.X + (size_t)chain * kFeat;
  if (mode) {
    for (int k = tid; k < FBH_PMAX * kFeatPos; k += kBlock) {
      const int v = k / kFeatPos, kk = k - v * kFeatPos;
      float g = 0.f;
      const int L = rgs[v] & 255, U = rgs[v] >> 8;
      const int cp = v < p ? feat_pos(L, U, kk) : -1;
      if (cp >= 0) {
        // left child's sums: the bins up to cp's own (bins are monotone in the rank, and a rank above cp lands in a
        // later bin: it has cp among the positions below it)
        const int sk = feat_bin(L, U, cp);
        long long w0 = 0, w1 = 0, t0 = 0, t1 = 0;
        for (int q = 0; q < kSegs; q++) {
          const long long h0 = (long long)H0[0][v][q], h1 = (long long)H1[0][v][q];
          t0 += h0;
          t1 += h1;
          if (q <= sk) {
            w0 += h0;
            w1 += h1;
          }
        }
        if (t0 > 0) {
          const double W = (double)t0 * (1.0 / kFxW), Sm = (double)t1 * (1.0 / kFxFeat);
          const double f = fmin(fmax((double)w0 / (double)t0, 1e-6), 1.0 - 1e-6);
          const double z = ((double)w1 * (1.0 / kFxFeat) - f * Sm) / __builtin_sqrt(W * f * (1.0 - f)) *
                           __builtin_sqrt(c.inv_s2);
          g = (float)(0.5 * z * z - 0.5 * log(f * (1.0 - f)));
        }
      }
      X[k] = g;
    }
    if (tid == 0) X[kFeat - 1] = (float)c.tree[0].d[nx] * 0.125f;
  } else {
    for (int k = tid; k < kFeat; k += kBlock) X[k] = 0.f;
  }
  if (tid < FBH_PMAX) A.rg[(size_t)chain * FBH_PMAX + tid] = rgs[tid];
  __syncthreads();
  img_copy(A.img + chain, &c, tid);
}

// ---- the exact split conditional at a birth node: the training target of the proposal network (no reference
// counterpart). Given the tree, sigma and the node's partial residuals, a birth's rule (v, c) has the conditional
//   q(v, c) proportional to (1 - PGl)(1 - PGr) PD / (U_v - L_v + 1) * exp(lh(left) + lh(right))
// over the rules with L_v <= c <= U_v whose children hold at least 5 rows each: BART's own rule probability times the
// factors of the birth ratio that depend on the rule (src/bart_kernel.cpp:441-451 and the likelihoods of :486); the
// remaining factors are the same for every rule. A network trained on labels drawn from q learns, in expectation,
// cross-entropy against q itself; its floor is the entropy of q, the uniform rule's level the log of the allowed
// class count. Sums of w and w r per (variable, cutpoint rank) are int64 fixed point at 2^-20 (order-free), and prefix
// sums along the ranks give every rule's left child at once.
constexpr int kTgtBlock = 256;
constexpr int kTgtGroup = 8;  // variables per histogram round
constexpr double kFxTgt = 1048576.0;

struct SplitNode {
  int n, p, ns;            // rows, variables, column stride of Q
  int d, ngv, ngood;       // depth and good-variable count of the node, good leaves of the tree
  double inv_s2, t2;
  const double* PG;        // depth prior table (:241-242)
  const uint8_t* Q;        // [p][ns] cutpoint ranks
  uint16_t rg[FBH_PMAX];   // L | U << 8 per variable (kNoRange: no cut)
};

struct SplitLds {
  unsigned long long h0[kTgtGroup][256], h1[kTgtGroup][256];
  unsigned hn[kTgtGroup][256];
  float lw[FBH_PMAX][256];   // log weight of each rule, -inf where the rule does not qualify
  double sv[FBH_PMAX], se[FBH_PMAX];
  double red[kTgtBlock / 64];
  double mx, Z, H;
};

__device__ __forceinline__ unsigned long long shfl_up_u64(unsigned long long x, int o) {
  const unsigned lo = __shfl_up((unsigned)x, o), hi = __shfl_up((unsigned)(x >> 32), o);
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ unsigned long long shfl_u64(unsigned long long x, int l) {
  const unsigned lo = __shfl((unsigned)x, l), hi = __shfl((unsigned)(x >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

// q at the node over all threads of a kTgtBlock workgroup. row(i, &wf, &wrf) says whether row i is in the node and
// gives w and w r at 2^-20. Returns false when no rule qualifies; else S.lw, S.mx, S.sv (per-variable totals of
// exp(lw - mx)), S.Z (their sum) and S.H (the entropy of q) are set.
template <class RowFn>
__device__ bool split_conditional(const SplitNode& N, SplitLds& S, RowFn row) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int k = tid; k < FBH_PMAX * 256; k += kTgtBlock) (&S.lw[0][0])[k] = -INFINITY;
  const double pb = 0.5;
  for (int g0 = 0; g0 < N.p; g0 += kTgtGroup) {
    for (int k = tid; k < kTgtGroup * 256; k += kTgtBlock) {
      (&S.h0[0][0])[k] = 0;
      (&S.h1[0][0])[k] = 0;
      (&S.hn[0][0])[k] = 0;
    }
    __syncthreads();
    const int ng = min(kTgtGroup, N.p - g0);
    for (int i = tid; i < N.n; i += kTgtBlock) {
      long long wf, wrf;
      if (!row(i, wf, wrf)) continue;
      for (int g = 0; g < ng; g++) {
        const int k = N.Q[(size_t)(g0 + g) * N.ns + i];
        atomicAdd(&S.h0[g][k], (unsigned long long)wf);
        atomicAdd(&S.h1[g][k], (unsigned long long)wrf);
        atomicAdd(&S.hn[g][k], 1u);
      }
    }
    __syncthreads();
    // one wave per variable: lane l holds ranks 4 l .. 4 l + 3, an inclusive scan over the lanes gives the left sums
    for (int g = wave; g < ng; g += kTgtBlock / 64) {
      const int v = g0 + g;
      const int L = N.rg[v] & 255, U = N.rg[v] >> 8;
      unsigned long long a0[4], a1[4];
      unsigned an[4];
      unsigned long long t0 = 0, t1 = 0;
      unsigned tn = 0;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        t0 += S.h0[g][4 * lane + e];
        t1 += S.h1[g][4 * lane + e];
        tn += S.hn[g][4 * lane + e];
        a0[e] = t0;
        a1[e] = t1;
        an[e] = tn;
      }
      unsigned long long i0 = t0, i1 = t1;
      unsigned in_ = tn;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y0 = shfl_up_u64(i0, o), y1 = shfl_up_u64(i1, o);
        const unsigned yn = __shfl_up(in_, o);
        if (lane >= o) {
          i0 += y0;
          i1 += y1;
          in_ += yn;
        }
      }
      const long long B0 = (long long)shfl_u64(i0, 63), B1 = (long long)shfl_u64(i1, 63);
      const int Nn = (int)__shfl(in_, 63);
      const unsigned long long x0 = i0 - t0, x1 = i1 - t1;
      const unsigned xn = in_ - tn;
      if (U < L) continue;
      const double PGd = N.PG[N.d + 1];
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int k = 4 * lane + e;
        const int cn = (int)(xn + an[e]);
        if (k < L || k > U || cn < 5 || Nn - cn < 5) continue;
        const long long c0 = (long long)(x0 + a0[e]), c1 = (long long)(x1 + a1[e]);
        double PGl, PGr;
        if (N.ngv > 1) {
          PGl = PGr = PGd;
        } else {
          PGl = (k - 1 < L) ? 0 : PGd;
          PGr = (U < k + 1) ? 0 : PGd;
        }
        const double PD = (N.ngood > 1) ? 1 - pb : ((PGr == 0 && PGl == 0) ? 1 : 1 - pb);
        const double s0 = (double)c0 * (1.0 / kFxTgt), s1 = (double)c1 * (1.0 / kFxTgt);
        const double r0 = (double)(B0 - c0) * (1.0 / kFxTgt), r1 = (double)(B1 - c1) * (1.0 / kFxTgt);
        const double lw = log((1 - PGl) * (1 - PGr) * PD / (double)(U - L + 1)) +
                          lh(s0 * N.inv_s2, s1 * N.inv_s2, N.t2) + lh(r0 * N.inv_s2, r1 * N.inv_s2, N.t2);
        S.lw[v][k] = (float)lw;
      }
    }
    __syncthreads();
  }
  // the largest log weight, then per-variable totals of exp(lw - mx) and of exp(lw - mx) (lw - mx) (fixed order)
  float m = -INFINITY;
  for (int k = tid; k < N.p * 256; k += kTgtBlock) m = fmaxf(m, (&S.lw[0][0])[k]);
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
  if (lane == 0) S.red[wave] = (double)m;
  __syncthreads();
  if (tid == 0) {
    double r = S.red[0];
    for (int w = 1; w < kTgtBlock / 64; w++) r = fmax(r, S.red[w]);
    S.mx = r;
  }
  __syncthreads();
  const double mx = S.mx;
  if (!(mx > -INFINITY)) return false;
  for (int v = wave; v < N.p; v += kTgtBlock / 64) {
    double s = 0, e = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float l = S.lw[v][lane + 64 * q];
      if (l > -INFINITY) {
        const double x = (double)l - mx, ex = exp(x);
        s += ex;
        e += ex * x;
      }
    }
    for (int o = 32; o > 0; o >>= 1) {
      s += __shfl_xor(s, o);
      e += __shfl_xor(e, o);
    }
    if (lane == 0) {
      S.sv[v] = s;
      S.se[v] = e;
    }
  }
  __syncthreads();
  if (tid == 0) {
    double Z = 0, E = 0;
    for (int v = 0; v < N.p; v++) {
      Z += S.sv[v];
      E += S.se[v];
    }
    S.Z = Z;
    S.H = log(Z) - E / Z;
  }
  __syncthreads();
  return true;
}

// inverse CDF of q at u in (0, 1), classes in the order v * 255 + c (one wave; every lane returns the class)
__device__ int split_draw(const SplitNode& N, const SplitLds& S, double u, int lane) {
  const double target = u * S.Z;
  double acc = 0;
  int v = -1, vlast = -1;
  for (int q = 0; q < N.p; q++) {
    if (!(S.sv[q] > 0)) continue;
    if (acc + S.sv[q] > target) {
      v = q;
      break;
    }
    acc += S.sv[q];
    vlast = q;
  }
  if (v < 0) {  // the total fell short of the target by rounding: the last variable with mass
    v = vlast;
    acc -= S.sv[vlast];
  }
  int cut = -1, clast = -1;
  for (int c0 = 0; c0 < 256; c0 += 64) {
    const float l = S.lw[v][c0 + lane];
    const double e = l > -INFINITY ? exp((double)l - S.mx) : 0.0;
    double pre = e;
    for (int o = 1; o < 64; o <<= 1) {
      const double y = __shfl_up(pre, o);
      if (lane >= o) pre += y;
    }
    const uint64_t hit = __ballot(e > 0 && acc + pre > target);
    const uint64_t has = __ballot(e > 0);
    if (has) clast = c0 + 63 - __builtin_clzll(has);
    if (hit) {
      cut = c0 + __builtin_ctzll(hit);
      break;
    }
    acc += __shfl(pre, 63);
  }
  if (cut < 0) cut = clast;
  return v * kCuts + cut;
}

__device__ __forceinline__ double mix_unif(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// after stage 1 of a collecting sweep with the conditional target: each chain whose move is a birth appends (features
// of the node, its ranges, a rule drawn from q, the entropy of q and the log of its allowed class count) to the pairs
__global__ __launch_bounds__(kTgtBlock) void k_lp_target(LpArgs A) {
  __shared__ SplitLds S;
  __shared__ SplitNode N;
  __shared__ int go, nx_s;
  __shared__ long long slot;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int chain = blockIdx.x, j = A.j;
  const KArgs& a = A.a;
  const Ctl& c = A.img[chain];
  if (tid == 0) {
    const int mode = c.cv[0], nx = c.cnx[0];
    const unsigned long long used = __atomic_load_n(A.pcount, __ATOMIC_RELAXED);
    const int it = A.st[chain].it;
    const unsigned long long key = ((unsigned long long)chain << 40) ^ ((unsigned long long)it * (unsigned)a.m + j);
    const bool take = A.keep >= 1.f || mix_unif(A.tseed ^ 0xD1B54A32D192ED03ull ^ (key * 0x9E3779B97F4A7C15ull)) <
                                           (double)A.keep;
    go = mode == 1 && take && (int)c.tree[0].d[nx] + 1 < kMaxDepth && (long long)used < (long long)A.pcap;
    nx_s = nx;
    N.n = c.cd.n;
    N.p = c.cd.p;
    N.ns = c.cd.ns;
    N.d = c.tree[0].d[nx];
    N.ngv = c.tree[0].ngv[nx];
    N.ngood = c.ngood[0];
    N.inv_s2 = c.inv_s2;
    N.t2 = (double)c.cd.tau * c.cd.tau;
    N.PG = a.PG;
    N.Q = a.q + c.cd.q_off;
  }
  if (tid < FBH_PMAX) N.rg[tid] = A.rg[(size_t)chain * FBH_PMAX + tid];
  __syncthreads();
  if (!go) return;
  const int nx = nx_s;
  const uint8_t* lf = a.leaf + c.cd.leaf_off + (size_t)j * c.cd.ns;
  const double* Y = a.Y + c.cd.row_off;
  const float* Wt = a.W + c.cd.row_off;
  const long long* AFX = A.AFX + c.cd.row_off;
  const long long thfx = fit_fx((double)c.tree[0].th[nx]);
  auto row = [&](int i, long long& wf, long long& wrf) -> bool {
    if (lf[i] != nx) return false;
    const double res = Y[i] - (double)(AFX[i] - thfx) * (1.0 / kFxFit);
    const double w = (double)Wt[i];
    wf = (long long)rint(w * kFxTgt);
    wrf = (long long)rint(w * res * kFxTgt);
    return true;
  };
  if (!split_conditional(N, S, row)) return;
  if (wave == 0) {
    const int it = A.st[chain].it;
    const double u = mix_unif(A.tseed + 0x9E3779B97F4A7C15ull * (((unsigned long long)chain << 40) ^
                                                                 ((unsigned long long)it * (unsigned)a.m + j) + 1));
    const int cls = split_draw(N, S, u, lane);
    if (lane == 0) {
      const unsigned long long k = atomicAdd(A.pcount, 1ull);
      slot = (long long)k < (long long)A.pcap ? (long long)k : -1;
      if (slot >= 0) {
        int allowed = 0;
        for (int v = 0; v < N.p; v++) {
          const int L = N.rg[v] & 255, U = N.rg[v] >> 8;
          if (U >= L) allowed += U - L + 1;
        }
        A.ptg[slot] = cls;
        A.pent[2 * slot] = (float)S.H;
        A.pent[2 * slot + 1] = (float)log((double)allowed);
      }
    }
  }
  __syncthreads();
  if (slot < 0) return;
  const float* X = A.X + (size_t)chain * kFeat;
  for (int q = tid; q < kFeat; q += kTgtBlock) A.px[(size_t)slot * kFeat + q] = X[q];
  if (tid < FBH_PMAX) A.prg[(size_t)slot * FBH_PMAX + tid] = N.rg[tid];
  if (A.pq) {
    float* q = A.pq + (size_t)slot * kClasses;
    for (int k = tid; k < kClasses; k += kTgtBlock) {
      const int v = k / kCuts, cc = k - v * kCuts;
      const float l = v < N.p ? S.lw[v][cc] : -INFINITY;
      q[k] = l > -INFINITY ? (float)(exp((double)l - S.mx) / S.Z) : 0.f;
    }
  }
}

// the split conditional of one node given on the host (the check of split_conditional / split_draw): q over the
// classes, the entropy, and the class drawn at each of nu uniforms
__global__ __launch_bounds__(kTgtBlock) void k_debug_split(SplitNode N0, const uint8_t* member, const double* w,
                                                          const double* r, const double* u, int nu, float* q_out,
                                                          int* cls_out, double* h_out) {
  __shared__ SplitLds S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto row = [&](int i, long long& wf, long long& wrf) -> bool {
    if (!member[i]) return false;
    wf = (long long)rint(w[i] * kFxTgt);
    wrf = (long long)rint(w[i] * r[i] * kFxTgt);
    return true;
  };
  const bool ok = split_conditional(N0, S, row);
  if (tid == 0) h_out[0] = ok ? S.H : -1.0;
  if (!ok) return;
  for (int k = tid; k < N0.p * 256; k += kTgtBlock) {
    const float l = (&S.lw[0][0])[k];
    q_out[k] = l > -INFINITY ? (float)(exp((double)l - S.mx) / S.Z) : 0.f;
  }
  for (int t = wave; t < nu; t += kTgtBlock / 64) {
    const int cls = split_draw(N0, S, u[t], lane);
    if (lane == 0) cls_out[t] = cls;
  }
}

// class k = v * 255 + c allowed at the node: v < p and L_v <= c <= U_v
__device__ __forceinline__ bool lp_valid(const uint16_t* rg, int k) {
  const int v = k / kCuts, cc = k - v * kCuts;
  const uint16_t r = rg[v];
  return cc >= (int)(r & 255) && cc <= (int)(r >> 8);
}

template <bool FX>
__global__ __attribute__((amdgpu_flat_work_group_size(kBlock, kBlock), amdgpu_waves_per_eu(4))) void k_lp_stage2(LpArgs A) {
  __shared__ Ctl c;
  __shared__ float red[kBlock];
  __shared__ float sv[FBH_PMAX];
  __shared__ uint16_t rg[FBH_PMAX];
  __shared__ float zmax, zsum;
  unsigned char* const dyn = fbh_dyn;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, dw = wave - 1;
  const int chain = blockIdx.x, j = A.j;
  const KArgs& a = A.a;
  img_copy(&c, A.img + chain, tid);
  if (tid < FBH_PMAX) {
    rg[tid] = A.rg[(size_t)chain * FBH_PMAX + tid];
    sv[tid] = 0.f;
  }
  __syncthreads();
  const RowPtrs P = lp_rows(a, c.cd);
  const int m = a.m, cap = a.node_cap, p = c.cd.p;
  TreeDev* TR = a.trees + c.cd.tree_off;
  LpState* S = A.st + chain;
  const LibmTables tb = a.tb;
  const int mode = c.cv[0], nx = c.cnx[0];
  const float alpha = A.alpha;
  // masked softmax of the chain's logits over the classes allowed at the node: max, total, per-variable totals
  if (mode && alpha > 0.f) {
    const float* z = A.Z + (size_t)chain * A.ldz;
    // one wave per variable, lanes over its allowed cuts (at most four per lane, loaded together)
    float mx = -INFINITY;
    for (int v = wave; v < p; v += kBlock / 64) {
      const int L = rg[v] & 255, U = rg[v] >> 8;
#pragma unroll
      for (int k = 0; k < (kCuts + 63) / 64; k++) {
        const int cc = L + lane + 64 * k;
        if (cc <= U) mx = fmaxf(mx, z[v * kCuts + cc]);
      }
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
    if (lane == 0) red[wave] = mx;
    __syncthreads();
    if (tid == 0) {
      float r = red[0];
      for (int w = 1; w < kBlock / 64; w++) r = fmaxf(r, red[w]);
      zmax = r;
    }
    __syncthreads();
    mx = zmax;
    // per-variable sums: one wave per variable group, lanes over the cuts (fixed order, so reproducible)
    for (int v = wave; v < p; v += kBlock / 64) {
      const int L = rg[v] & 255, U = rg[v] >> 8;
      float zc[(kCuts + 63) / 64];
#pragma unroll
      for (int k = 0; k < (kCuts + 63) / 64; k++) {
        const int cc = lane + 64 * k;
        zc[k] = (cc >= L && cc <= U) ? z[v * kCuts + cc] : 0.f;
      }
      float s = 0.f;
#pragma unroll
      for (int k = 0; k < (kCuts + 63) / 64; k++) {
        const int cc = lane + 64 * k;
        if (cc >= L && cc <= U) s += expf(zc[k] - mx);
      }
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) sv[v] = s;
    }
    __syncthreads();
    if (tid == 0) {
      float tot = 0.f;
      for (int v = 0; v < p; v++) tot += sv[v];
      zsum = tot;
    }
    __syncthreads();
  }
  Rng g{0, 0, 0, 0};
  if (wave == 0) {
    g = Rng{ufl(S->rng[0]), ufl(S->rng[1]), ufl(S->rng[2]), ufl(S->rng[3])};
    const TreeDev& t = c.tree[0];
    const double pb = 0.5;
    const int ngood = c.ngood[0], nnogs = c.nnogs[0];
    const uint64_t leafm = t.leaf, used = t.used;
    const double PBx = ngood == 0 ? 0 : (used == 1ull ? 1 : pb);
    const float* z = A.Z + (size_t)chain * A.ldz;
    int cmode = 0, cv = 0, cc = 0;
    double lpr = 0;
    uint64_t cand = 0;
    if (mode == 1) {
      // The rule: uniform with probability 1 - alpha (:435-437), else from the network. The mixture uniform is drawn
      // only when alpha > 0, so that with alpha = 0 the generator stream is the plain proposal's, draw for draw
      // (the cut is drawn, as in :436-437, only when the range is not empty and the node is above the depth limit).
      int v = -1, cut = -1;
      const bool deep_ok = (int)t.d[nx] + 1 < kMaxDepth;
      if (!(alpha > 0.f) || unif_d(g) >= (double)alpha) {
        v = (int)(unif_d(g) * p);
        if (v >= p) v = p - 1;
        const uint16_t r = rg[v];
        const int L = r & 255, U = r >> 8;
        if (U >= L && deep_ok) {
          cut = L + (int)(unif_d(g) * (U - L + 1));
          if (cut > U) cut = U;
        }
      } else {
        // inverse CDF of the masked softmax with a 48-bit uniform: variable by its total, then the cut
        const uint32_t h1 = rng_next(g) >> 8, h2 = rng_next(g) >> 8;
        const double u = ((double)h1 * 16777216.0 + (double)h2 + 0.5) * (1.0 / 281474976710656.0);
        const double target = u * (double)zsum;
        double acc = 0;
        int vlast = -1;
        for (int q = 0; q < p; q++) {
          if (!(sv[q] > 0.f)) continue;
          if (acc + (double)sv[q] > target) {
            v = q;
            break;
          }
          acc += (double)sv[q];
          vlast = q;
        }
        if (v < 0 && vlast >= 0) {  // the total fell short of the target by rounding: the last variable with mass
          v = vlast;
          acc -= (double)sv[vlast];
        }
        const uint16_t r = v >= 0 ? rg[v] : kNoRange;  // no variable with mass (non-finite logits): no birth
        const int L = r & 255, U = r >> 8;
        cut = U >= L ? U : -1;
        for (int c0 = L; c0 <= U; c0 += 64) {
          const int ci = c0 + lane;
          const float e = ci <= U ? expf(z[v * kCuts + ci] - zmax) : 0.f;
          // inclusive prefix over the lanes
          float pre = e;
          for (int o = 1; o < 64; o <<= 1) {
            const float y = __shfl_up(pre, o);
            if (lane >= o) pre += y;
          }
          const uint64_t hit = __ballot(ci <= U && acc + (double)pre > target);
          if (hit) {
            cut = c0 + __builtin_ctzll(hit);
            break;
          }
          acc += (double)__shfl(pre, 63);
        }
      }
      if (v >= 0 && cut >= 0 && deep_ok) {
        const uint16_t r = rg[v];
        const int bL = r & 255, bU = r >> 8;
        const int dnx = t.d[nx];
        const double PGnx = c.PG[dnx];
        double PGly, PGry;
        if (t.ngv[nx] > 1) {
          PGly = PGry = c.PG[dnx + 1];
        } else {
          PGly = (cut - 1 < bL) ? 0 : c.PG[dnx + 1];
          PGry = (bU < cut + 1) ? 0 : c.PG[dnx + 1];
        }
        const double PDy = (ngood > 1) ? 1 - pb : ((PGry == 0 && PGly == 0) ? 1 : 1 - pb);
        double Pnogy;
        if (nx == 0) {
          Pnogy = 1;
        } else {
          const int pa = t.par[nx];
          const bool panog = BITM(leafm, t.l[pa] & 63) && BITM(leafm, t.r[pa] & 63);
          Pnogy = panog ? 1.0 / nnogs : 1.0 / (nnogs + 1);
        }
        const double Pbotx = 1.0 / ngood;
        const double pr = (PGnx * (1 - PGly) * (1 - PGry) * PDy * Pnogy) / ((1 - PGnx) * Pbotx * PBx);
        double fac = 1.0;
        if (alpha > 0.f) {
          const double qu = (1.0 / p) / (double)(bU - bL + 1);
          const double qn = (double)expf(z[v * kCuts + cut] - zmax) / (double)zsum;
          fac = qu / ((1.0 - (double)alpha) * qu + (double)alpha * qn);
        }
        lpr = log(pr * fac);
        cmode = 1;
        cv = v;
        cc = cut;
        const int L = first_free(used, -1, cap), R = L < 0 ? -1 : first_free(used, L, cap);
        if (L >= 0 && R >= 0) cand = (1ull << L) | (1ull << R);
      }
    } else if (mode == 2) {
      const int dny = t.d[nx], l = t.l[nx], r = t.r[nx];
      const double PGny = c.PG[dny];
      const double PGlx = t.ngv[l] > 0 ? c.PG[dny + 1] : 0, PGrx = t.ngv[r] > 0 ? c.PG[dny + 1] : 0;
      const double PBy = (nx == 0) ? 1 : pb;
      const int ng2 = ngood - (t.ngv[l] > 0) - (t.ngv[r] > 0) + 1;
      const double Pboty = 1.0 / ng2, PDx = 1 - PBx, Pnogx = 1.0 / nnogs;
      const double pr = ((1 - PGny) * PBy * Pboty) / (PGny * (1 - PGlx) * (1 - PGrx) * PDx * Pnogx);
      double fac = 1.0;
      if (alpha > 0.f) {
        const int v = t.v[nx], cut = t.c[nx];
        const uint16_t rr = rg[v];
        const int bL = rr & 255, bU = rr >> 8;
        const double qu = (1.0 / p) / (double)(bU - bL + 1);
        const double qn = (double)expf(z[v * kCuts + cut] - zmax) / (double)zsum;
        fac = ((1.0 - (double)alpha) * qu + (double)alpha * qn) / qu;
      }
      lpr = log(pr * fac);
      cmode = 2;
    }
    if (lane == 0) {
      c.cmode[0] = cmode;
      c.cnx[0] = nx;
      c.cv[0] = cv;
      c.cc[0] = cc;
      c.lpr[0] = lpr;
      c.cand[0] = cand;
      c.nL = c.nR = 0;
    }
  }
  __syncthreads();
  const int cmode = c.cmode[0];
  if (cmode == 1 && wave > 0) {
    TileIn<kRowsStreamL> T;
    data_pass<kRowsStreamL, false, FX, true, false>(c, dyn, P, dw, lane, j - 1, j, 0, false, true, false, T, false);
  }
  __syncthreads();
  if (wave == 0) {
    if (cmode == 1) ctl_combine<FX>(c, 0, false, 1, lane);
    TreeDev& t = c.tree[0];
    const double inv_s2 = c.inv_s2;
    const double t2 = (double)c.cd.tau * c.cd.tau;
    const double a_prior = 1 / t2;
    const double lu = log_unif(g, tb);
    int n_birth = S->n_birth, n_death = S->n_death, n_cap = S->n_cap;
    int nmode = 0, nL_ = 0, nR_ = 0, nl_ = 0, nr_ = 0, acc_birth = 0;
    if (cmode == 1) {
      if (c.nL >= 5 && c.nR >= 5) {
        const double bl = c.sL0 * inv_s2, Ml = c.sL1 * inv_s2, br = c.sR0 * inv_s2, Mr = c.sR1 * inv_s2;
        double lalpha = c.lpr[0] + lh(bl, Ml, t2) + lh(br, Mr, t2) - lh(bl + br, Ml + Mr, t2);
        if (lalpha > 0) lalpha = 0;
        if (lu < lalpha) {
          const int L = first_free(t.used, -1, cap);
          const int Rr = L < 0 ? -1 : first_free(t.used, L, cap);
          if (L < 0 || Rr < 0) {
            n_cap++;
          } else {
            if (lane == 0) {
              t.v[nx] = (uint8_t)c.cv[0];
              t.c[nx] = (uint8_t)c.cc[0];
              t.l[nx] = (uint8_t)L;
              t.r[nx] = (uint8_t)Rr;
              t.par[L] = t.par[Rr] = (uint8_t)nx;
              t.l[L] = t.r[L] = t.l[Rr] = t.r[Rr] = 0;
              t.d[L] = t.d[Rr] = (uint8_t)(t.d[nx] + 1);
              t.used |= (1ull << L) | (1ull << Rr);
              t.leaf = (t.leaf & ~(1ull << nx)) | (1ull << L) | (1ull << Rr);
              c.S0[L] = c.sL0;
              c.S1[L] = c.sL1;
              c.S0[Rr] = c.sR0;
              c.S1[Rr] = c.sR1;
            }
            wave_sync();
            const int gL = count_goodvars_wave(t, L, c.cd.nc, p, lane);
            const int gR = count_goodvars_wave(t, Rr, c.cd.nc, p, lane);
            if (lane == 0) {
              t.ngv[L] = (uint8_t)gL;
              t.ngv[Rr] = (uint8_t)gR;
            }
            nmode = 1;
            nL_ = L;
            nR_ = Rr;
            n_birth++;
            acc_birth = 1;
          }
        }
      }
    } else if (cmode == 2) {
      const int l = t.l[nx], r = t.r[nx];
      const double bl = c.S0[l] * inv_s2, Ml = c.S1[l] * inv_s2, br = c.S0[r] * inv_s2, Mr = c.S1[r] * inv_s2;
      double lalpha = c.lpr[0] + lh(bl + br, Ml + Mr, t2) - lh(bl, Ml, t2) - lh(br, Mr, t2);
      if (lalpha > 0) lalpha = 0;
      if (lu < lalpha) {
        const double m0 = c.S0[l] + c.S0[r], m1 = c.S1[l] + c.S1[r];
        wave_sync();
        if (lane == 0) {
          c.S0[nx] = m0;
          c.S1[nx] = m1;
          t.used &= ~((1ull << l) | (1ull << r));
          t.leaf = (t.leaf & ~((1ull << l) | (1ull << r))) | (1ull << nx);
          t.l[nx] = t.r[nx] = t.v[nx] = t.c[nx] = 0;
        }
        nmode = 2;
        nl_ = l;
        nr_ = r;
        n_death++;
      }
    }
    wave_sync();
    // one normal per slot (:520-528)
    float zn = 0.f;
    for (int s = 0; s < cap; s++) {
      const uint32_t x1 = rng_next(g);
      const uint32_t x2 = rng_next(g);
      if (lane == s) zn = normal_f(x1, x2, tb);
    }
    const int s = lane;
    c.tho[s] = c.thr[0][s].th;
    if (s < cap && BITM(t.leaf, s) && BITM(t.used, s)) {
      const double b = c.S0[s] * inv_s2, M = c.S1[s] * inv_s2;
      t.th[s] = (float)(M / (a_prior + b) + (double)zn / __builtin_sqrt(a_prior + b));
    }
    wave_sync();
    c.thn[s] = (double)t.th[s];
    const int it = S->it;
    if (it >= a.nskip && a.keep_trees)
      a.trees_out[c.cd.tout_off + ((size_t)(it - a.nskip) * m + j) * kNodes + lane] = pack_word(t, lane);
    copy_tree(&TR[j], &t, lane);
    // an accepted birth's (features, ranges, class) becomes a training pair (with the conditional target, k_lp_target
    // has taken every proposed birth's node already)
    if (acc_birth && A.pcap > 0 && A.target == 0) {
      unsigned long long k = 0;
      if (lane == 0) k = atomicAdd(A.pcount, 1ull);
      k = __shfl(k, 0);
      if ((long)k < A.pcap) {
        const float* X = A.X + (size_t)chain * kFeat;
        for (int q = lane; q < kFeat; q += 64) A.px[(size_t)k * kFeat + q] = X[q];
        if (lane < FBH_PMAX) A.prg[(size_t)k * FBH_PMAX + lane] = rg[lane];
        if (lane == 0) A.ptg[k] = c.cv[0] * kCuts + c.cc[0];
      }
    }
    if (lane == 0) {
      c.pmode = nmode;
      c.pnx = nx;
      c.pv = c.cv[0];
      c.pc = c.cc[0];
      c.pL = nL_;
      c.pR = nR_;
      c.pl = nl_;
      c.pr = nr_;
      S->rng[0] = g.s0;
      S->rng[1] = g.s1;
      S->rng[2] = g.s2;
      S->rng[3] = g.s3;
      S->n_birth = n_birth;
      S->n_death = n_death;
      S->n_cap = n_cap;
    }
  }
  __syncthreads();
  img_copy(A.img + chain, &c, tid);
}

// closing pass of a sweep (apply tree m-1, residual sum of squares, posterior-mean sum) and the sigma draw (:536-558)
template <bool FX>
__global__ __launch_bounds__(kBlock) void k_lp_close(LpArgs A) {
  __shared__ Ctl c;
  unsigned char* const dyn = fbh_dyn;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, dw = wave - 1;
  const int chain = blockIdx.x;
  const KArgs& a = A.a;
  img_copy(&c, A.img + chain, tid);
  __syncthreads();
  const RowPtrs P = lp_rows(a, c.cd);
  const int m = a.m;
  LpState* S = A.st + chain;
  const int it = S->it;
  const bool keep = it >= a.nskip;
  if (wave > 0) {
    lp_afx_apply(c, P, A.AFX + c.cd.row_off, m - 1, dw, lane);
    TileIn<kRowsStreamL> T;
    data_pass<kRowsStreamL, false, FX, true, true>(c, dyn, P, dw, lane, m - 1, m, 0, true, false, keep, T, false);
  }
  __syncthreads();
  if (wave == 0) {
    ctl_combine<FX>(c, 0, true, 0, lane);
    Rng g{ufl(S->rng[0]), ufl(S->rng[1]), ufl(S->rng[2]), ufl(S->rng[3])};
    const float nul = c.cd.nu * c.cd.lambda;
    const float nnu = (float)c.cd.n_total + c.cd.nu;
    const double sigma = __builtin_sqrt(((double)nul + c.rss) / (2.0 * rgamma_dev(g, (double)nnu / 2.0, a.tb)));
    if (lane == 0) {
      a.sigma_out[c.cd.sig_off + it] = (float)sigma;
      S->sigma = sigma;
      S->it = it + 1;
      S->rng[0] = g.s0;
      S->rng[1] = g.s1;
      S->rng[2] = g.s2;
      S->rng[3] = g.s3;
      c.pmode = 0;
    }
  }
  __syncthreads();
  img_copy(A.img + chain, &c, tid);
}

// control images and fixed-point fits from the batch state (all trees applied, between sweeps)
__global__ __launch_bounds__(kBlock) void k_lp_begin(LpArgs A) {
  __shared__ Ctl c;
  const int tid = threadIdx.x, lane = tid & 63;
  const int chain = blockIdx.x;
  const KArgs& a = A.a;
  if (tid < 64) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(a.desc + chain);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&c.cd);
    for (int w = lane; w < (int)(sizeof(ChainDesc) / 4); w += 64) dst[w] = src[w];
    c.PG[lane] = a.PG[lane];
    if (lane < 2) c.rot[lane] = lane ? 21u : 11u;
    if (lane == 0) c.pmode = 0;
  }
  const ChainState* cs = a.state + chain;
  LpState* S = A.st + chain;
  if (tid == 0) {
    for (int w = 0; w < 4; w++) S->rng[w] = cs->rng[w];
    S->sigma = cs->sigma;
    S->it = cs->it;
    S->mode = S->nx = 0;
    S->n_birth = cs->n_birth;
    S->n_death = cs->n_death;
    S->n_cap = cs->n_cap;
  }
  __syncthreads();
  const int n = c.cd.n, m = a.m;
  const TreeDev* TR = a.trees + c.cd.tree_off;
  const uint8_t* LEAF = a.leaf + c.cd.leaf_off;
  long long* AFX = A.AFX + c.cd.row_off;
  for (int i = tid; i < n; i += kBlock) {
    long long s = 0;
    for (int jj = 0; jj < m; jj++) s += fit_fx((double)TR[jj].th[LEAF[(size_t)jj * c.cd.ns + i]]);
    AFX[i] = s;
  }
  __syncthreads();
  img_copy(A.img + chain, &c, tid);
}

// a minibatch of training pairs into the network's batch buffers: features, class, allowed classes from the ranges
__global__ void k_lp_gather(const float* __restrict__ px, const uint16_t* __restrict__ prg, const int* __restrict__ ptg,
                            const float* __restrict__ pq, const int* __restrict__ idx, int M, float* __restrict__ x,
                            int* __restrict__ tgt, uint8_t* __restrict__ mask, float* __restrict__ soft) {
  const int b = blockIdx.x, tid = threadIdx.x;
  if (b >= M) return;
  const long r = idx[b];
  for (int k = tid; k < kFeat; k += blockDim.x) x[(size_t)b * kFeat + k] = px[(size_t)r * kFeat + k];
  for (int k = tid; k < kClasses; k += blockDim.x) mask[(size_t)b * kClasses + k] = lp_valid(prg + r * FBH_PMAX, k);
  if (pq)
    for (int k = tid; k < kClasses; k += blockDim.x) soft[(size_t)b * kClasses + k] = pq[(size_t)r * kClasses + k];
  if (tid == 0) tgt[b] = ptg[r];
}

// the learned runs' scalars back into the batch state
__global__ void k_lp_end(LpArgs A, int C) {
  const int chain = blockIdx.x * blockDim.x + threadIdx.x;
  if (chain >= C) return;
  const LpState* S = A.st + chain;
  ChainState* cs = A.a.state + chain;
  for (int w = 0; w < 4; w++) cs->rng[w] = S->rng[w];
  cs->sigma = S->sigma;
  cs->it = S->it;
  cs->n_birth = S->n_birth;
  cs->n_death = S->n_death;
  cs->n_cap = S->n_cap;
}


}  // namespace fbh
// ==================================================================== host side
using namespace fbh;

struct fbh_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  // C device chains (workgroups); CL logical chains, each W consecutive members (W = 1 without a row split);
  // logical chain c has rows [row0[c*W + k], row0[c*W + k + 1]) in member k
  int C = 0, CL = 0, W = 1;
  std::vector<int> row0;
  unsigned long long* d_xbuf = nullptr;
  uint32_t* d_xctr = nullptr;
  int m = 0, nskip = 0, ndpost = 0, probit = 0, keep_trees = 0, stats_mode = 0, node_cap = kNodes;
  int sweeps_per_launch = 25;
  int it = 0;
  std::vector<ChainDesc> hdesc;
  ChainDesc* d_desc = nullptr;
  ChainState* d_state = nullptr;
  TreeDev* d_trees = nullptr;
  uint8_t* d_q = nullptr;
  double* d_Y = nullptr;
  float* d_W = nullptr;
  float* d_yin = nullptr;
  double* d_AF = nullptr;
  double* d_TM = nullptr;
  uint8_t* d_ybin = nullptr;
  uint8_t* d_leaf = nullptr;
  float* d_sigma = nullptr;
  uint64_t* d_tout = nullptr;
  double* d_PG = nullptr;
  float* d_trm = nullptr;
  TraceRec* d_trace = nullptr;
  int trace_cap = 0;
  Variant var{};
  size_t dyn_bytes = 0;
  int af_off = 0;
  // the kernel variants this batch may switch between at launch boundaries (the state between launches is the same):
  // [dense][resident]; fits = all workgroups resident at once (required by a row split)
  Variant vars[2][kKinds]{};
  bool kind_ok[kKinds]{};
  bool fits[2][kKinds]{};
  // the packing: workgroup w runs device chain order[w]
  int packing = 0, ncu = 0;
  std::vector<int> order;
  int* d_order = nullptr;
  // launches captured in a graph per fbh_batch_run call (released by the next fbh_batch_sync or by destroy)
  int graph = 0;
  std::vector<hipGraph_t> graphs;
  std::vector<hipGraphExec_t> gexecs;
  int nmax = 0;
  LibmTables tb{};
  size_t rows_total = 0, q_total = 0, leaf_total = 0;
  // timing: a start/stop event pair around each plain launch, or around each launch of a captured graph (nl launches)
  struct Timed {
    hipEvent_t e0, e1;
    int nl;
  };
  std::vector<Timed> ev;
  double bytes_per_sweep = 0;
  // learned birth proposal
  fbh_mlp* net = nullptr;
  float alpha = 0.f;
  Ctl* d_img = nullptr;
  LpState* d_lps = nullptr;
  long long* d_afx = nullptr;
  float* d_X = nullptr;
  uint16_t* d_rg = nullptr;
  float* d_px = nullptr;
  uint16_t* d_prg = nullptr;
  int* d_ptg = nullptr;
  unsigned long long* d_pcount = nullptr;
  int* d_idx = nullptr;
  float* d_pent = nullptr;
  float* d_pq = nullptr;
  long pcap = 0, hold = 0;
  int target = 0;
  float keep = 1.f;
  unsigned long long tseed = 0;
  std::vector<hipEvent_t> lev;  // start/stop pairs per learned sweep
  // probit latent redraw: 1 in the sweep kernel, 2 by the k_pr_* kernels after every sweep (buffers made on first use)
  int pr_mode = 1;
  uint32_t pr_K = 0;
  int pr_nwmax = 0;
  uint4* d_jump = nullptr;
  uint2* d_slots = nullptr;
  uint4* d_ck = nullptr;
  float2* d_ev = nullptr;
  PrChain* d_pc = nullptr;
  double2* d_btot = nullptr;
  float* d_E = nullptr;
  int* d_lo = nullptr;
  uint8_t* d_nw = nullptr;
  unsigned long long* d_mask = nullptr;
  uint32_t* d_acc = nullptr;
  unsigned* d_spot = nullptr;
  int* d_tab = nullptr;
  uint32_t* d_entry = nullptr;
  int* d_gtab = nullptr;
  uint32_t* d_gentry = nullptr;
};

static void free_batch(fbh_batch* b) {
  if (!b) return;
  if (b->device >= 0) (void)hipSetDevice(b->device);
  void* ptrs[] = {b->d_desc, b->d_state, b->d_trees, b->d_q,   b->d_Y,   b->d_W,   b->d_yin,    b->d_AF,
                  b->d_TM,   b->d_ybin,  b->d_leaf,  b->d_sigma, b->d_tout, b->d_PG, b->d_trm, b->d_trace,
                  b->d_img,  b->d_lps,   b->d_afx,   b->d_X,     b->d_rg,   b->d_px, b->d_prg, b->d_ptg,
                  b->d_pcount, b->d_idx, b->d_xbuf, b->d_xctr, b->d_order, b->d_pent, b->d_pq,
                  b->d_jump,  b->d_slots, b->d_ck,    b->d_E,     b->d_lo,   b->d_nw, b->d_mask, b->d_acc,
                  b->d_spot,  b->d_ev,    b->d_tab,   b->d_entry, b->d_btot, b->d_gtab, b->d_gentry, b->d_pc};
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto g : b->gexecs) (void)hipGraphExecDestroy(g);
  for (auto g : b->graphs) (void)hipGraphDestroy(g);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& t : b->ev) {
    (void)hipEventDestroy(t.e0);
    (void)hipEventDestroy(t.e1);
  }
  for (auto e : b->lev) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

template <class T>
static int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T)) != hipSuccess) {
    set_error(FBH_ENOMEM, "hipMalloc failed");
    return FBH_ENOMEM;
  }
  return FBH_OK;
}

// ---- probit redraw spread over the device (k_pr_*)
constexpr int kPrAutoChains = 16;  // automatic choice: the k_pr_* kernels for batches of at most this many chains
// the generator's state transition applied to a host state (rng_step)
static void host_xo_step(uint32_t s[4]) {
  const uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = (s[3] << 11) | (s[3] >> 21);
}
// out[k][j] = T^(128 2^k) e_j: 128 steps from each unit state, then repeated squaring (column j of M M is M applied to
// column j of M)
static void host_jump_tables(std::vector<uint32_t>& out) {
  out.assign((size_t)kJumpLevels * 128 * 4, 0u);
  for (int j = 0; j < 128; j++) {
    uint32_t st[4] = {0u, 0u, 0u, 0u};
    st[j >> 5] = 1u << (j & 31);
    for (int k = 0; k < 2 * kPrSlots; k++) host_xo_step(st);
    std::memcpy(&out[(size_t)j * 4], st, 16);
  }
  for (int L = 1; L < kJumpLevels; L++) {
    const uint32_t* P = &out[(size_t)(L - 1) * 512];
    for (int j = 0; j < 128; j++) {
      uint32_t o[4] = {0u, 0u, 0u, 0u};
      for (int bit = 0; bit < 128; bit++)
        if ((P[4 * j + (bit >> 5)] >> (bit & 31)) & 1u)
          for (int w = 0; w < 4; w++) o[w] ^= P[4 * bit + w];
      std::memcpy(&out[((size_t)L * 128 + j) * 4], o, 16);
    }
  }
}
// the state 64 blocks slots (128 blocks generator steps) after s, through the tables
static void host_jump(const std::vector<uint32_t>& J, const uint32_t s[4], uint64_t blocks, uint32_t out[4]) {
  uint32_t st[4] = {s[0], s[1], s[2], s[3]};
  for (int k = 0; k < kJumpLevels; k++) {
    if (!((blocks >> k) & 1u)) continue;
    uint32_t o[4] = {0u, 0u, 0u, 0u};
    for (int bit = 0; bit < 128; bit++)
      if ((st[bit >> 5] >> (bit & 31)) & 1u)
        for (int w = 0; w < 4; w++) o[w] ^= J[((size_t)k * 128 + bit) * 4 + w];
    std::memcpy(st, o, 16);
  }
  std::memcpy(out, st, 16);
}

// rows of the longest logical chain (a split chain's members together)
static int pr_nmax(const fbh_batch* b) {
  int m = 1;
  for (int c = 0; c < b->CL; c++) m = std::max(m, b->hdesc[(size_t)c * b->W].n_total);
  return m;
}
static int pr_blocks(const fbh_batch* b) { return (pr_nmax(b) + 63) / 64; }
static int pr_groups(const fbh_batch* b) { return (pr_blocks(b) + kPrGroup - 1) / kPrGroup; }
static size_t pr_table_lds(int nwmax) { return (size_t)64 * nwmax * 8 + 64 * 8; }

// buffers of the device-wide redraw; K = 0: room for twice the longest chain's rows and its widest band (an attempt
// rejects with probability at most 1/2, so a chain uses about n + E_n <= 2 n slots); a smaller K (a multiple of 64)
// sends the slots past it to the walk's own stepping
static int pr_alloc(fbh_batch* b, uint32_t K) {
  const int nm = pr_nmax(b);
  const int nwmax = std::min(255, (int)std::ceil((7.0 * std::sqrt(2.0 * nm) + 64.0) / 64.0) + 1);
  const uint64_t Kd = ((2ull * (uint64_t)nm + 64ull * (uint64_t)nwmax + 4096ull) + 63ull) / 64ull * 64ull;
  if (K == 0) K = (uint32_t)Kd;
  if (K % kPrSlots || K < (uint32_t)kPrSlots || (uint64_t)K / kPrSlots > (1ull << kJumpLevels)) {
    set_error(FBH_EINVAL, "probit redraw: slots must be a positive multiple of 64 below 2^26");
    return FBH_EINVAL;
  }
  int rc = FBH_OK;
  if (!b->d_jump) {
    std::vector<uint32_t> J;
    host_jump_tables(J);
    if ((rc = dalloc(&b->d_jump, (size_t)kJumpLevels * 128))) return rc;
    if (hipMemcpy(b->d_jump, J.data(), J.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      set_error(FBH_EHIP, "probit redraw: hipMemcpy failed");
      return FBH_EHIP;
    }
  }
  if (!b->d_E) {
    const size_t ro = b->rows_total;
    if ((rc = dalloc(&b->d_pc, (size_t)b->CL)) || (rc = dalloc(&b->d_ev, ro)) ||
        (rc = dalloc(&b->d_btot, (size_t)b->C * ((nm + 255) / 256 + 1))) || (rc = dalloc(&b->d_E, ro)) ||
        (rc = dalloc(&b->d_lo, ro)) || (rc = dalloc(&b->d_nw, ro)) ||
        (rc = dalloc(&b->d_mask, ro * (size_t)nwmax)) || (rc = dalloc(&b->d_acc, ro)) ||
        (rc = dalloc(&b->d_tab, (size_t)b->C * pr_blocks(b) * 64 * (size_t)nwmax)) ||
        (rc = dalloc(&b->d_entry, (size_t)b->C * pr_blocks(b))) ||
        (rc = dalloc(&b->d_gtab, (size_t)b->C * pr_groups(b) * 64 * (size_t)nwmax)) ||
        (rc = dalloc(&b->d_gentry, (size_t)b->C * pr_groups(b))) ||
        (rc = dalloc(&b->d_spot, (size_t)b->C)))
      return rc;
    if (hipMemset(b->d_spot, 0, sizeof(unsigned) * b->C) != hipSuccess) {
      set_error(FBH_EHIP, "probit redraw: hipMemset failed");
      return FBH_EHIP;
    }
    b->pr_nwmax = nwmax;
    // the logical chains: member 0's pool position; a split chain's members hold `share` rows each at a stride of
    // share rounded up to 64 (the pools keep each member's rows 64-row aligned)
    std::vector<PrChain> hpc((size_t)b->CL);
    for (int c = 0; c < b->CL; c++) {
      const ChainDesc& d0 = b->hdesc[(size_t)c * b->W];
      PrChain& q = hpc[c];
      q.row_off = d0.row_off;
      q.n = d0.n_total;
      q.share = b->W > 1 ? d0.n : d0.n_total;
      q.share_pad = (q.share + 63) / 64 * 64;
      q.W = b->W;
      q.dc0 = c * b->W;
      q.offset = d0.offset;
    }
    if (hipMemcpy(b->d_pc, hpc.data(), sizeof(PrChain) * b->CL, hipMemcpyHostToDevice) != hipSuccess) {
      set_error(FBH_EHIP, "probit redraw: hipMemcpy failed");
      return FBH_EHIP;
    }
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(k_pr_table), hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)pr_table_lds(nwmax)) != hipSuccess) {
      set_error(FBH_EHIP, "probit redraw: hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
      return FBH_EHIP;
    }
  }
  if (K != b->pr_K) {
    (void)hipFree(b->d_slots);
    (void)hipFree(b->d_ck);
    b->d_slots = nullptr;
    b->d_ck = nullptr;
    b->pr_K = 0;
    if ((rc = dalloc(&b->d_slots, (size_t)b->C * K)) || (rc = dalloc(&b->d_ck, (size_t)b->C * (K / kPrSlots))))
      return rc;
    b->pr_K = K;
  }
  return FBH_OK;
}

static PrArgs make_pr_args(const fbh_batch* b) {
  PrArgs A;
  A.pc = b->d_pc;
  A.state = b->d_state;
  A.AF = b->d_AF;
  A.ybin = b->d_ybin;
  A.Y = b->d_Y;
  A.tb = b->tb;
  A.jump = b->d_jump;
  A.slots = b->d_slots;
  A.ck = b->d_ck;
  A.ev = b->d_ev;
  A.btot = b->d_btot;
  A.nb256 = (pr_nmax(b) + 255) / 256;
  A.E = b->d_E;
  A.lo = b->d_lo;
  A.nw = b->d_nw;
  A.mask = b->d_mask;
  A.acc = b->d_acc;
  A.spot = b->d_spot;
  A.K = b->pr_K;
  A.nwmax = b->pr_nwmax;
  A.tab = b->d_tab;
  A.entry = b->d_entry;
  A.nbmax = pr_blocks(b);
  A.gtab = b->d_gtab;
  A.gentry = b->d_gentry;
  A.ngmax = pr_groups(b);
  A.wmax = 64 * b->pr_nwmax;
  return A;
}

// one sweep's redraw for every chain, on the batch's stream (after the sweep kernel has left the fit in HBM)
static hipError_t pr_launch(const fbh_batch* b) {
  const PrArgs A = make_pr_args(b);
  const unsigned C = (unsigned)b->CL, nm = (unsigned)pr_nmax(b);
  hipLaunchKernelGGL(k_pr_rows, dim3((nm + 255) / 256, C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_band, dim3(C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_bandrows, dim3((nm + 255) / 256, C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_gen, dim3((A.K / kPrSlots + 63) / 64, C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_decide, dim3((nm + 3) / 4, C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_table, dim3((unsigned)A.nbmax, C), dim3(256), pr_table_lds(A.nwmax), b->stream, A);
  hipLaunchKernelGGL(k_pr_compose, dim3((unsigned)A.ngmax, C), dim3(256), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_chain, dim3(C), dim3(64), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_expand, dim3((unsigned)A.ngmax, C), dim3(64), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_walk, dim3((unsigned)A.nbmax, C), dim3(64), 0, b->stream, A);
  hipLaunchKernelGGL(k_pr_value, dim3((nm + 255) / 256, C), dim3(256), 0, b->stream, A);
  return hipGetLastError();
}

static KArgs make_args(const fbh_batch* b) {
  KArgs a;
  a.desc = b->d_desc;
  a.state = b->d_state;
  a.trees = b->d_trees;
  a.q = b->d_q;
  a.Y = b->d_Y;
  a.yf = b->d_yin;
  a.W = b->d_W;
  a.AF = b->d_AF;
  a.TM = b->d_TM;
  a.ybin = b->d_ybin;
  a.leaf = b->d_leaf;
  a.sigma_out = b->d_sigma;
  a.trees_out = b->d_tout;
  a.PG = b->d_PG;
  a.tb = b->tb;
  a.trace = b->d_trace;
  a.trace_cap = b->trace_cap;
  a.m = b->m;
  a.nskip = b->nskip;
  a.ndpost = b->ndpost;
  a.probit = b->probit;
  a.keep_trees = b->keep_trees;
  a.stats_mode = b->stats_mode;
  a.node_cap = b->node_cap;
  a.it0 = a.it1 = 0;
  a.af_off = b->af_off;
  a.xbuf = b->d_xbuf;
  a.xctr = b->d_xctr;
  a.order = b->d_order;
  a.pr_ext = 0;
  return a;
}

// The packing (workgroup -> device chain). 0: the batch order. 1: serpentine -- device chains by rows, largest first,
// dealt to workgroups in rounds of one per CU with every second round reversed, so that where the dispatcher gives
// workgroups w and w + ncu the same CU, a large chain shares its CU with a small one. 2: largest first (longest
// processing time first, for batches with more chains than resident slots). Chains are independent, so the packing
// changes timing only, never results.
static void packing_order(const fbh_batch* b, int mode, std::vector<int>& ord) {
  const int C = b->C;
  ord.resize(C);
  for (int w = 0; w < C; w++) ord[w] = w;
  if (mode == 0) return;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return b->hdesc[x].n > b->hdesc[y].n; });
  if (mode == 1) {
    const int r = std::max(1, b->ncu);
    for (int s0 = r, k = 1; s0 < C; s0 += r, k++)
      if (k & 1) std::reverse(ord.begin() + s0, ord.begin() + std::min(C, s0 + r));
  }
}

static size_t dyn_of(const fbh_batch* b, int kind) {
  return align16(kind_scratch(kind)) + (kind_resident(kind) ? 8 * afl_doubles(((size_t)b->nmax + 15) / 16 * 16) : 0);
}

static void use_variant(fbh_batch* b, int kind, bool dense) {
  b->var = b->vars[dense][kind];
  b->af_off = (int)align16(b->var.scratch);
  b->dyn_bytes = dyn_of(b, kind);
}

// the tile kind for tile_rows (0, 1024, 1536) and resident (0 automatic, 1 streamed, 2 in LDS); -1 if none fits.
// Automatic: the running fit in LDS when it fits -- 1024-row tiles for the dense build, 1536 otherwise -- except that
// tile_rows = 1024 alone keeps the streamed kind.
static int pick_kind(const fbh_batch* b, int tile_rows, int resident, bool dense) {
  if (resident == 1) return tile_rows == 1536 ? -1 : 0;
  if (tile_rows == 1536) return b->kind_ok[2] ? 2 : -1;
  if (tile_rows == 1024) return resident == 2 ? (b->kind_ok[1] ? 1 : -1) : 0;
  const int pref[2] = {dense ? 1 : 2, dense ? 2 : 1};
  for (int k : pref)
    if (b->kind_ok[k]) return k;
  return resident == 2 ? -1 : 0;
}

// switch the packing, the tile kind and the build between launches (waits for the batch's queued work)
static int configure(fbh_batch* b, int packing, int kind, bool dense) {
  if (packing < 0 || packing > 2) {
    set_error(FBH_EINVAL, "fbh_batch_configure: packing must be 0, 1 or 2");
    return FBH_EINVAL;
  }
  if (kind < 0 || !b->kind_ok[kind]) {
    set_error(FBH_EINVAL, "fbh_batch_configure: no such tile configuration for this batch (a running fit in LDS "
                          "needs the longest chain's rows to fit beside the tile, 1536-row tiles need it in LDS)");
    return FBH_EINVAL;
  }
  if (b->W > 1 && !b->fits[dense][kind]) {
    set_error(FBH_EINVAL, "fbh_batch_configure: the row split's workgroups would not all be resident at once");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  packing_order(b, packing, b->order);
  FBH_HIP_CHECK(hipMemcpy(b->d_order, b->order.data(), sizeof(int) * b->C, hipMemcpyHostToDevice));
  b->packing = packing;
  use_variant(b, kind, dense);
  return FBH_OK;
}

constexpr int kSplitRows = 3072;  // the automatic row split's fewest rows per member

// rows per member of a row split: ceil(n / W); for probit chains rounded up to a multiple of 64, so that the latent
// redraw's blocks of 64 rows never straddle two members (0 if the last member would then be empty)
static int split_share(int n, int W, bool probit) {
  int s = (n + W - 1) / W;
  if (probit) {
    s = (s + 63) / 64 * 64;
    if ((long)(W - 1) * s >= n) return 0;
  }
  return s;
}

extern "C" int fbh_batch_create(const fbh_chain_in* chains, int nchains, const fbh_opts* o, fbh_batch** out) {
  if (!out || !chains || !o || nchains < 1) {
    set_error(FBH_EINVAL, "fbh_batch_create: null argument or no chains");
    return FBH_EINVAL;
  }
  *out = nullptr;
  if (o->m < 1 || o->m > FBH_MMAX || o->nskip < 0 || o->ndpost < 0 || !o->pgtab ||
      (o->stats_mode != FBH_STATS_REF16 && o->stats_mode != FBH_STATS_FX64) ||
      (o->node_cap != 0 && (o->node_cap < 3 || o->node_cap > kNodes)) ||
      (o->block_threads != 0 && o->block_threads != kBlock) ||
      (o->tile_rows != 0 && o->tile_rows != Geo<kRowsStream>::kTile && o->tile_rows != Geo<kRowsRes>::kTile) ||
      o->packing < 0 || o->packing > 2 || o->chains_per_cu < 0 || o->chains_per_cu > 2 || o->resident < 0 ||
      o->resident > 2 || o->probit_redraw < 0 || o->probit_redraw > 2) {
    set_error(FBH_EINVAL, "fbh_batch_create: bad options (m, nskip, ndpost, pgtab, stats_mode, node_cap, "
                          "block_threads, tile_rows, packing, chains_per_cu, resident or probit_redraw)");
    return FBH_EINVAL;
  }
  int W = o->row_split > 1 ? o->row_split : 1;
  const bool pb = o->probit != 0;
  if (o->row_split < 0 && o->stats_mode == FBH_STATS_FX64 && !(pb && o->probit_redraw == 1)) {
    // automatic: as many workgroups per chain as leave one workgroup per CU, with at least kSplitRows rows per
    // member (fewer, and the exchange after every pass costs more than the shorter pass saves: measured on one
    // chain, 4,096 wbart rows run fastest on one workgroup, 24,576 probit rows on 8)
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, o->device) != hipSuccess) ncu = 0;
    W = std::max(1, std::min(64, ncu / nchains));
    for (int c = 0; c < nchains; c++) {
      W = std::max(1, std::min(W, chains[c].n / kSplitRows));
      while (W > 1 && !split_share(chains[c].n, W, pb)) W--;
    }
  }
  if (W > 1 && (o->stats_mode != FBH_STATS_FX64 || W > 64 || (pb && o->probit_redraw == 1))) {
    set_error(FBH_EINVAL, "fbh_batch_create: row_split needs FX64 statistics (at most 64; for probit chains the "
                          "latent redraw spread over the device)");
    return FBH_EINVAL;
  }
  for (int c = 0; W > 1 && c < nchains; c++)
    if (chains[c].n < 64 * W || !split_share(chains[c].n, W, pb)) {
      set_error(FBH_EINVAL, "fbh_batch_create: row_split " + std::to_string(W) + " needs n >= " +
                                std::to_string(64 * W) + " in every chain (probit: rows for every member in "
                                "multiples of 64)");
      return FBH_EINVAL;
    }
  for (int c = 0; c < nchains; c++) {
    const fbh_chain_in& ci = chains[c];
    if (ci.n < 1 || ci.n > FBH_NMAX || ci.p < 1 || ci.p > FBH_PMAX || !ci.q || !ci.wt || !ci.numcut ||
        (!o->probit && !ci.y) || (o->probit && !ci.ybin)) {
      set_error(FBH_EINVAL, "fbh_batch_create: chain " + std::to_string(c) + " has a bad shape or null input");
      return FBH_EINVAL;
    }
    for (int v = 0; v < ci.p; v++)
      if (ci.numcut[v] < 0 || ci.numcut[v] > 255) {
        set_error(FBH_EINVAL, "fbh_batch_create: numcut outside 0..255");
        return FBH_EINVAL;
      }
    double ymax = 0.0, wmax = 0.0;
    for (int i = 0; i < ci.n; i++) {
      if (!std::isfinite(ci.wt[i]) || (!o->probit && !std::isfinite(ci.y[i]))) {
        set_error(FBH_EINVAL, "fbh_batch_create: non-finite y or weight in chain " + std::to_string(c));
        return FBH_EINVAL;
      }
      wmax = std::max(wmax, std::fabs((double)ci.wt[i]));
      if (!o->probit) ymax = std::max(ymax, std::fabs((double)ci.y[i]));
    }
    if (o->stats_mode == FBH_STATS_FX64) {
      // Fixed-point headroom: per-row images of w and w r at 2^32 and of w e^2 at 2^24 are summed in int64, so
      // |sum w r| < 2^31 and sum w e^2 < 2^39 must hold at every pass. With residuals bounded by 4 max|y| (probit
      // latents: by 64), these bounds keep a factor of 2 of margin; past them the sums could wrap.
      const double rb = o->probit ? 64.0 : 4.0 * ymax;
      const double s1 = (double)ci.n * wmax * rb, s2 = (double)ci.n * wmax * rb * rb;
      if (!(s1 < 1073741824.0) || !(s2 < 274877906944.0)) {
        set_error(FBH_EINVAL, "fbh_batch_create: chain " + std::to_string(c) +
                                  " is outside the FX64 range (n max(wt) (4 max|y|)^2 must stay below 2^38 and "
                                  "n max(wt) 4 max|y| below 2^30): rescale y or use REF16 statistics");
        return FBH_EINVAL;
      }
    }
  }
  int rc = ensure_device(o->device);
  if (rc) return rc;
  fbh_batch* b = new fbh_batch();
  b->device = o->device;
  b->CL = nchains;
  b->W = W;
  b->C = nchains * W;
  b->m = o->m;
  b->nskip = o->nskip;
  b->ndpost = o->ndpost;
  b->probit = o->probit ? 1 : 0;
  // a split chain's rows are spread over its members: only the device-wide redraw sees them all
  b->pr_mode = W > 1 ? 2 : (o->probit_redraw ? o->probit_redraw : (b->C <= kPrAutoChains ? 2 : 1));
  b->keep_trees = o->keep_trees ? 1 : 0;
  b->stats_mode = o->stats_mode;
  b->node_cap = o->node_cap ? o->node_cap : kNodes;
  b->sweeps_per_launch = o->sweeps_per_launch > 0 ? o->sweeps_per_launch : 25;
  if ((rc = libm_tables(b->device, &b->tb))) {
    free_batch(b);
    return rc;
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error(FBH_EHIP, "hipStreamCreate failed");
    free_batch(b);
    return FBH_EHIP;
  }
  // ---- pools: one device chain per (logical chain, member); a member's rows are a contiguous range of the chain's
  const int nchains_l = nchains;
  nchains = b->C;
  b->hdesc.resize(nchains);
  b->row0.assign((size_t)nchains + 1, 0);
  size_t qo = 0, ro = 0, lo = 0;
  const int total = b->nskip + b->ndpost;
  for (int dc = 0; dc < nchains; dc++) {
    const int c = dc / W, k = dc % W;
    const fbh_chain_in& ci = chains[c];
    const int share = W > 1 ? split_share(ci.n, W, b->probit != 0) : ci.n;
    const int r0 = std::min(ci.n, k * share), r1 = std::min(ci.n, (k + 1) * share);
    b->row0[dc] = r0;
    if (k == W - 1) b->row0[dc + 1] = ci.n;
    ChainDesc& d = b->hdesc[dc];
    std::memset(&d, 0, sizeof d);
    d.n = r1 - r0;
    d.ns = (d.n + 15) / 16 * 16;
    d.n_total = ci.n;
    d.split = W;
    d.member = k;
    d.group = c;
    d.keep = k == 0;
    d.p = ci.p;
    d.tau = ci.tau;
    d.lambda = ci.lambda;
    d.nu = ci.nu;
    d.sigest = ci.sigest;
    d.offset = ci.offset;
    b->nmax = std::max(b->nmax, d.n);
    int rn = 0;
    for (int v = 0; v < FBH_PMAX; v++) {
      d.nc[v] = v < ci.p ? ci.numcut[v] : 0;
      if (v < ci.p && d.nc[v] >= 1) rn++;
    }
    d.root_ngv = rn;
    d.q_off = qo;
    d.row_off = ro;
    d.leaf_off = lo;
    d.tree_off = (uint64_t)dc * b->m;
    d.sig_off = (uint64_t)dc * total;
    d.tout_off = (uint64_t)c * b->ndpost * b->m * kNodes;  // only member 0 writes the kept trees
    qo += (size_t)ci.p * d.ns;
    ro += ((size_t)d.n + 63) / 64 * 64;  // keep each chain's rows 512-byte aligned
    lo += ((size_t)b->m * d.ns + 255) / 256 * 256;
    b->bytes_per_sweep += 37.0 * d.n * (b->m + 1);
  }
  b->q_total = qo;
  b->rows_total = ro;
  b->leaf_total = lo;
  // ---- kernel variant and LDS: the running fit stays resident in LDS when control state + tile scratch + 8 bytes a row
  // fit the per-workgroup budget that lets two chains share a CU (160 KiB of LDS per CU on gfx950); otherwise it streams.
  {
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, b->device) != hipSuccess)
      lds_max = 65536;
    const size_t budget = std::min<size_t>((size_t)lds_max, kLdsBudgetPerWG);
    for (int k = 0; k < kKinds; k++)
      b->kind_ok[k] = !kind_resident(k) ||
                      kCtlBytes + dyn_of(b, k) <= budget;
    if (hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, b->device) != hipSuccess) b->ncu = 0;
    const bool fx = b->stats_mode == FBH_STATS_FX64;
    for (int d = 0; d < 2; d++)
      for (int k = 0; k < kKinds; k++) {
        Variant& v = b->vars[d][k];
        v = pick_variant(k, fx, b->probit != 0, d != 0);
        if (!b->kind_ok[k]) continue;
        const size_t dyn = dyn_of(b, k);
        if (kCtlBytes + dyn > (size_t)lds_max) {
          set_error(FBH_EINVAL, "fbh_batch_create: LDS need exceeds the device limit");
          free_batch(b);
          return FBH_EINVAL;
        }
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(v.fn), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)dyn) != hipSuccess) {
          set_error(FBH_EHIP, "fbh_batch_create: hipFuncSetAttribute(MaxDynamicSharedMemorySize) failed");
          free_batch(b);
          return FBH_EHIP;
        }
        // workgroups resident per CU: the occupancy calculator's figure, capped by the dispatcher's reservation of
        // two waves per SIMD for a 5-wave workgroup (one workgroup per CU at 168 VGPRs, two at 128)
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(v.fn), kBlock, dyn) !=
            hipSuccess)
          per_cu = 0;
        per_cu = std::min(per_cu, d ? 2 : 1);
        b->fits[d][k] = (long)per_cu * b->ncu >= (long)nchains;
      }
    // automatic: the dense build when there are more workgroups than CUs -- except for probit batches, whose dense
    // variants spill heavily (the autotuner still tries them)
    bool dense = o->chains_per_cu == 2 || (o->chains_per_cu == 0 && nchains > b->ncu && !o->probit);
    int kind = pick_kind(b, o->tile_rows, o->resident, dense);
    if (kind < 0) {
      set_error(FBH_EINVAL, "fbh_batch_create: tile_rows / resident ask for the running fit in LDS, and the longest "
                            "chain's rows do not fit there beside the tile");
      free_batch(b);
      return FBH_EINVAL;
    }
    if (W > 1) {
      // every member must be resident at once: the exchange waits for all of them. An automatic choice of build
      // takes the other build if only that one fits.
      if (!b->fits[dense][kind] && o->chains_per_cu == 0 && b->fits[!dense][kind]) dense = !dense;
      if (!b->fits[dense][kind]) {
        set_error(FBH_EINVAL, "fbh_batch_create: " + std::to_string(nchains) + " workgroups of a row split do not "
                                  "fit on the device at once (" + std::to_string(dense ? 2 : 1) + " per CU x " +
                                  std::to_string(b->ncu) + " CUs)");
        free_batch(b);
        return FBH_EINVAL;
      }
    }
    use_variant(b, kind, dense);
  }
  const size_t tout = b->keep_trees ? (size_t)nchains_l * b->ndpost * b->m * kNodes : 1;
  if (W > 1) {
    if ((rc = dalloc(&b->d_xbuf, (size_t)nchains_l * 2 * W * kXw)) || (rc = dalloc(&b->d_xctr, (size_t)nchains_l)) ||
        hipMemset(b->d_xctr, 0, sizeof(uint32_t) * nchains_l) != hipSuccess) {
      free_batch(b);
      return rc ? rc : FBH_EHIP;
    }
  }
  if ((rc = dalloc(&b->d_desc, nchains)) || (rc = dalloc(&b->d_state, nchains)) ||
      (rc = dalloc(&b->d_trees, (size_t)nchains * b->m)) || (rc = dalloc(&b->d_q, qo)) ||
      (rc = dalloc(&b->d_Y, ro)) || (rc = dalloc(&b->d_W, ro)) || (rc = dalloc(&b->d_yin, ro)) ||
      (rc = dalloc(&b->d_AF, ro)) || (rc = dalloc(&b->d_TM, ro)) || (rc = dalloc(&b->d_ybin, ro)) ||
      (rc = dalloc(&b->d_leaf, lo)) || (rc = dalloc(&b->d_sigma, (size_t)nchains * std::max(total, 1))) ||
      (rc = dalloc(&b->d_tout, tout)) || (rc = dalloc(&b->d_PG, kNodes)) || (rc = dalloc(&b->d_trm, ro)) ||
      (rc = dalloc(&b->d_order, nchains))) {
    free_batch(b);
    return rc;
  }
  b->packing = o->packing;
  b->graph = o->graph ? 1 : 0;
  packing_order(b, b->packing, b->order);
  // ---- staging on the host, one copy per pool
  {
    std::vector<uint8_t> hq(qo), hyb(ro, 0);
    std::vector<float> hw(ro, 0.f), hy(ro, 0.f);
    for (int dc = 0; dc < nchains; dc++) {
      const fbh_chain_in& ci = chains[dc / W];
      const ChainDesc& d = b->hdesc[dc];
      const int r0 = b->row0[dc];
      for (int v = 0; v < ci.p; v++)
        std::memcpy(hq.data() + d.q_off + (size_t)v * d.ns, ci.q + (size_t)v * ci.n + r0, (size_t)d.n);
      std::memcpy(hw.data() + d.row_off, ci.wt + r0, sizeof(float) * d.n);
      if (!o->probit) std::memcpy(hy.data() + d.row_off, ci.y + r0, sizeof(float) * d.n);
      if (o->probit)
        for (int i = 0; i < d.n; i++) hyb[d.row_off + i] = ci.ybin[r0 + i] != 0;
    }
    std::vector<double> pg(kNodes);
    for (int d = 0; d < kNodes; d++) pg[d] = (double)o->pgtab[d];
    std::vector<unsigned> seeds(nchains);
    for (int dc = 0; dc < nchains; dc++) seeds[dc] = chains[dc / W].seed;  // members share the generator
    unsigned* d_seeds = nullptr;
    if (hipMalloc(&d_seeds, sizeof(unsigned) * nchains) != hipSuccess) {
      set_error(FBH_ENOMEM, "hipMalloc failed");
      free_batch(b);
      return FBH_ENOMEM;
    }
    hipError_t e = hipSuccess;
    auto cp = [&](void* dst, const void* src, size_t bytes) {
      if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, b->stream);
    };
    cp(b->d_desc, b->hdesc.data(), sizeof(ChainDesc) * nchains);
    cp(b->d_q, hq.data(), qo);
    cp(b->d_W, hw.data(), sizeof(float) * ro);
    cp(b->d_yin, hy.data(), sizeof(float) * ro);
    cp(b->d_ybin, hyb.data(), ro);
    cp(b->d_PG, pg.data(), sizeof(double) * kNodes);
    cp(d_seeds, seeds.data(), sizeof(unsigned) * nchains);
    cp(b->d_order, b->order.data(), sizeof(int) * nchains);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_leaf, 0, lo, b->stream);
    // rows past n in the row pools are read (never used) by the last quad of rows of a tile: keep them zero
    if (e == hipSuccess) e = hipMemsetAsync(b->d_AF, 0, sizeof(double) * ro, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->d_Y, 0, sizeof(double) * ro, b->stream);
    // the initial probit latents (:384-386) are a redraw with the running fit at zero: batches that redraw with the
    // k_pr_* kernels draw them there too
    const bool par = b->probit && b->pr_mode == 2;
    if (e == hipSuccess) {
      KArgs a = make_args(b);
      a.pr_ext = par ? 1 : 0;
      hipLaunchKernelGGL(k_init, dim3(nchains), dim3(256), 0, b->stream, a, d_seeds);
      e = hipGetLastError();
    }
    if (e == hipSuccess && par) {
      const int rc = pr_alloc(b, 0);
      if (rc) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(d_seeds);
        free_batch(b);
        return rc;
      }
      e = pr_launch(b);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    (void)hipFree(d_seeds);
    if (e != hipSuccess) {
      set_error(FBH_EHIP, std::string("fbh_batch_create: ") + hipGetErrorString(e));
      free_batch(b);
      return FBH_EHIP;
    }
  }
  *out = b;
  return FBH_OK;
}

extern "C" int fbh_batch_run(fbh_batch* b, int nsweeps) {
  if (!b || nsweeps < 0) {
    set_error(FBH_EINVAL, "fbh_batch_run: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  const int total = b->nskip + b->ndpost;
  const int end = std::min(total, b->it + nsweeps);
  if (b->it >= end) return FBH_OK;
  // the launches of this call: plain, each between a pair of timing events; or, with graph = 1, captured into one
  // graph of kernel nodes in stream order that is launched once between a pair of timing events. With the probit
  // redraw spread over the device a launch is one sweep, followed by the k_pr_* kernels.
  const bool par = b->probit && b->pr_mode == 2;
  if (par) {
    const int rc = pr_alloc(b, b->pr_K);
    if (rc) return rc;
  }
  const int per = par ? 1 : b->sweeps_per_launch;
  const int nl = (end - b->it + per - 1) / per;
  const int npair = b->graph ? 1 : nl;
  const size_t ev0 = b->ev.size();
  for (int k = 0; k < npair; k++) {
    fbh_batch::Timed t{nullptr, nullptr, b->graph ? nl : 1};
    FBH_HIP_CHECK(hipEventCreate(&t.e0));
    if (hipEventCreate(&t.e1) != hipSuccess) {
      (void)hipEventDestroy(t.e0);
      set_error(FBH_EHIP, "fbh_batch_run: hipEventCreate failed");
      return FBH_EHIP;
    }
    b->ev.push_back(t);
  }
  hipError_t e = hipSuccess;
  hipGraph_t g = nullptr;
  if (b->graph) e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
  KArgs a = make_args(b);
  a.pr_ext = par ? 1 : 0;
  for (int k = 0, it = b->it; k < nl && e == hipSuccess; k++) {
    a.it0 = it;
    a.it1 = std::min(end, it + per);
    void* args[] = {&a};
    if (!b->graph) e = hipEventRecord(b->ev[ev0 + k].e0, b->stream);
    if (e == hipSuccess)
      e = hipLaunchKernel(reinterpret_cast<const void*>(b->var.fn), dim3(b->C), dim3(kBlock), args, b->dyn_bytes,
                          b->stream);
    if (e == hipSuccess && par) e = pr_launch(b);
    if (e == hipSuccess && !b->graph) e = hipEventRecord(b->ev[ev0 + k].e1, b->stream);
    it = a.it1;
  }
  if (b->graph) {
    const hipError_t ec = hipStreamEndCapture(b->stream, &g);
    if (e == hipSuccess) e = ec;
    if (g) b->graphs.push_back(g);
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (x) b->gexecs.push_back(x);
    if (e == hipSuccess) e = hipEventRecord(b->ev[ev0].e0, b->stream);
    if (e == hipSuccess) e = hipGraphLaunch(x, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->ev[ev0].e1, b->stream);
  }
  if (e != hipSuccess) {
    set_error(FBH_EHIP, std::string("fbh_batch_run: ") + hipGetErrorString(e));
    return FBH_EHIP;
  }
  b->it = end;
  return FBH_OK;
}

extern "C" int fbh_batch_configure(fbh_batch* b, int packing, int tile_rows, int resident, int chains_per_cu) {
  if (!b || (tile_rows != 0 && tile_rows != Geo<kRowsStream>::kTile && tile_rows != Geo<kRowsRes>::kTile) ||
      resident < 0 || resident > 2 || chains_per_cu < 0 || chains_per_cu > 2) {
    set_error(FBH_EINVAL, "fbh_batch_configure: bad argument");
    return FBH_EINVAL;
  }
  if (b->net) {
    set_error(FBH_EINVAL, "fbh_batch_configure: not for a batch set up for the learned proposal");
    return FBH_EINVAL;
  }
  const bool dense = chains_per_cu == 0 ? b->var.dense : chains_per_cu == 2;
  return configure(b, packing, pick_kind(b, tile_rows, resident, dense), dense);
}

extern "C" int fbh_batch_probit_redraw(fbh_batch* b, int mode, int slots) {
  if (!b || !b->probit || mode < 0 || mode > 2 || slots < 0) {
    set_error(FBH_EINVAL, "fbh_batch_probit_redraw: bad argument (a probit batch, mode 0..2, slots >= 0)");
    return FBH_EINVAL;
  }
  if (mode == 1 && b->W > 1) {
    set_error(FBH_EINVAL, "fbh_batch_probit_redraw: a chain split over workgroups redraws with the device-wide "
                          "kernels only");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  const int m = b->W > 1 ? 2 : (mode ? mode : (b->C <= kPrAutoChains ? 2 : 1));
  if (m == 2) {
    FBH_HIP_CHECK(hipStreamSynchronize(b->stream));  // the slot buffers may be replaced
    const int rc = pr_alloc(b, (uint32_t)slots);
    if (rc) return rc;
  }
  b->pr_mode = m;
  return FBH_OK;
}

extern "C" int fbh_batch_probit_info(fbh_batch* b, int* mode, int* slots, unsigned long long* spot) {
  if (!b) {
    set_error(FBH_EINVAL, "fbh_batch_probit_info: null batch");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  if (mode) *mode = b->probit ? b->pr_mode : 0;
  if (slots) *slots = (int)b->pr_K;
  if (spot) {
    *spot = 0;
    if (b->d_spot) {
      std::vector<unsigned> h((size_t)b->C);
      FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
      FBH_HIP_CHECK(hipMemcpy(h.data(), b->d_spot, sizeof(unsigned) * b->C, hipMemcpyDeviceToHost));
      for (unsigned v : h) *spot += v;
    }
  }
  return FBH_OK;
}

extern "C" int fbh_debug_jump(const uint32_t* state, unsigned long long blocks, uint32_t* out) {
  if (!state || !out || blocks >= (1ull << kJumpLevels)) {
    set_error(FBH_EINVAL, "fbh_debug_jump: bad argument");
    return FBH_EINVAL;
  }
  static const std::vector<uint32_t> J = [] {
    std::vector<uint32_t> v;
    host_jump_tables(v);
    return v;
  }();
  host_jump(J, state, blocks, out);
  return FBH_OK;
}

extern "C" int fbh_batch_set_order(fbh_batch* b, const int* order) {
  if (!b || !order) {
    set_error(FBH_EINVAL, "fbh_batch_set_order: null argument");
    return FBH_EINVAL;
  }
  std::vector<int> seen(b->CL, 0);
  for (int w = 0; w < b->CL; w++) {
    if (order[w] < 0 || order[w] >= b->CL || seen[order[w]]++) {
      set_error(FBH_EINVAL, "fbh_batch_set_order: not a permutation of the chains");
      return FBH_EINVAL;
    }
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  // logical chain order[w] takes workgroups w W .. w W + W - 1 (its row-split members, in member order)
  for (int w = 0; w < b->CL; w++)
    for (int k = 0; k < b->W; k++) b->order[(size_t)w * b->W + k] = order[w] * b->W + k;
  FBH_HIP_CHECK(hipMemcpy(b->d_order, b->order.data(), sizeof(int) * b->C, hipMemcpyHostToDevice));
  b->packing = -1;
  return FBH_OK;
}

extern "C" int fbh_batch_config(const fbh_batch* b, int* packing, int* tile_rows, int* resident, int* chains_per_cu,
                                int* graph) {
  if (!b) {
    set_error(FBH_EINVAL, "fbh_batch_config: null batch");
    return FBH_EINVAL;
  }
  if (packing) *packing = b->packing;
  if (tile_rows) *tile_rows = 256 * kind_rows(b->var.kind);
  if (resident) *resident = b->var.resident ? 2 : 1;
  if (chains_per_cu) *chains_per_cu = b->var.dense ? 2 : 1;
  if (graph) *graph = b->graph;
  return FBH_OK;
}

// Autotuner: times candidate launch configurations on the batch's own next sweeps -- real sweeps of every chain, whose
// results no candidate changes -- and keeps the fastest. Two stages: the build (one or two chains per CU; two only
// with more workgroups than CUs) x the tile kind (1024-row tiles with the running fit streamed or in LDS, 1536-row
// tiles with it in LDS), at largest-first packing; then the packing with the stage's winner.
// One untimed warm-up launch, then per stage two rounds of one launch of trial_sweeps sweeps per candidate, the
// second round in reverse order so that drift in the cost of a sweep (trees growing during burn-in) falls on every
// candidate alike.
extern "C" int fbh_batch_autotune(fbh_batch* b, int trial_sweeps, fbh_tune* out) {
  if (!b || trial_sweeps < 1) {
    set_error(FBH_EINVAL, "fbh_batch_autotune: bad argument");
    return FBH_EINVAL;
  }
  if (b->net) {
    set_error(FBH_EINVAL, "fbh_batch_autotune: not for a batch set up for the learned proposal");
    return FBH_EINVAL;
  }
  struct Cand {
    int packing, kind;
    bool dense;
    double ms;
  };
  std::vector<Cand> s1;
  for (int d = (b->C > b->ncu) ? 1 : 0; d >= 0; d--)
    for (int k = kKinds - 1; k >= 0; k--)
      if (b->kind_ok[k] && (b->W <= 1 || b->fits[d][k])) s1.push_back(Cand{2, k, d != 0, 0.0});
  if (s1.empty()) s1.push_back(Cand{2, b->var.kind, b->var.dense, 0.0});
  const int need = (1 + 2 * (int)s1.size() + 2 * 3) * trial_sweeps;
  if (b->it + need > b->nskip + b->ndpost) {
    set_error(FBH_EINVAL, "fbh_batch_autotune: " + std::to_string(need) + " sweeps needed, " +
                              std::to_string(b->nskip + b->ndpost - b->it) + " left");
    return FBH_EINVAL;
  }
  const int spl = b->sweeps_per_launch, g = b->graph;
  b->sweeps_per_launch = trial_sweeps;
  b->graph = 0;
  auto stage = [&](std::vector<Cand>& cs) -> int {
    for (int round = 0; round < 2; round++)
      for (size_t k = 0; k < cs.size(); k++) {
        Cand& cd = cs[round ? cs.size() - 1 - k : k];
        int rc = configure(b, cd.packing, cd.kind, cd.dense);
        if (!rc) rc = fbh_batch_run(b, trial_sweeps);
        if (!rc) rc = fbh_batch_sync(b);
        if (rc) return rc;
        float t = 0.f;
        if (hipEventElapsedTime(&t, b->ev.back().e0, b->ev.back().e1) != hipSuccess) {
          (void)hipGetLastError();
          set_error(FBH_EHIP, "fbh_batch_autotune: hipEventElapsedTime failed");
          return FBH_EHIP;
        }
        cd.ms += t / (2.0 * trial_sweeps);
      }
    return FBH_OK;
  };
  auto best_of = [](const std::vector<Cand>& cs) {
    size_t k = 0;
    for (size_t i = 1; i < cs.size(); i++)
      if (cs[i].ms < cs[k].ms) k = i;
    return cs[k];
  };
  int rc = fbh_batch_run(b, trial_sweeps);
  std::vector<Cand> s2;
  if (!rc) rc = stage(s1);
  if (!rc) {
    const Cand w = best_of(s1);
    for (int pk = 0; pk < 3; pk++) s2.push_back(Cand{pk, w.kind, w.dense, 0.0});
    rc = stage(s2);
  }
  b->sweeps_per_launch = spl;
  b->graph = g;
  if (rc) return rc;
  const Cand w = best_of(s2);
  if ((rc = configure(b, w.packing, w.kind, w.dense))) return rc;
  if (out) {
    std::memset(out, 0, sizeof *out);
    out->packing = w.packing;
    out->tile_rows = 256 * kind_rows(w.kind);
    out->resident = kind_resident(w.kind) ? 2 : 1;
    out->chains_per_cu = w.dense ? 2 : 1;
    out->sweeps = need;
    std::vector<Cand> all(s1);
    all.insert(all.end(), s2.begin(), s2.end());
    out->ncand = std::min<int>((int)all.size(), FBH_TUNE_MAX);
    for (int i = 0; i < out->ncand; i++) {
      out->cand_packing[i] = all[i].packing;
      out->cand_tile_rows[i] = 256 * kind_rows(all[i].kind);
      out->cand_resident[i] = kind_resident(all[i].kind) ? 2 : 1;
      out->cand_chains_per_cu[i] = all[i].dense ? 2 : 1;
      out->ms_per_sweep[i] = all[i].ms;
    }
  }
  return FBH_OK;
}

// ------------------------------------------------------------------ autotuner cache
// A text file of one line per tuned batch shape, "<key>\t<packing> <tile_rows> <resident> <chains_per_cu> <ms>",
// where ms is the chosen configuration's time per sweep. The key names what the time of a sweep depends on: the
// device (architecture string and CU count), the library build, statistics mode, probit, row split, chain count, trees,
// node cap, the rows per chain (minimum, quartiles and maximum, rounded up to 256-row steps) and the predictors per
// chain (range). The last line for a key wins, so a re-tune appended later replaces an earlier entry.
static std::string tune_key(const fbh_batch* b) {
  std::string arch = "unknown";
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, b->device) == hipSuccess) arch = pr.gcnArchName;
  else (void)hipGetLastError();
  std::vector<int> ns(b->CL), ps(b->CL);
  for (int c = 0; c < b->CL; c++) {
    const ChainDesc& d = b->hdesc[(size_t)c * b->W];
    ns[c] = (d.n_total + 255) / 256 * 256;
    ps[c] = d.p;
  }
  std::sort(ns.begin(), ns.end());
  std::sort(ps.begin(), ps.end());
  auto q = [&](int k) { return ns[(size_t)k * (ns.size() - 1) / 4]; };
  std::string build = __DATE__ "_" __TIME__;
  for (char& ch : build)
    if (ch == ' ') ch = '_';
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "fbh1 arch=%s cu=%d build=%s stats=%s probit=%d split=%d chains=%d ntree=%d cap=%d rows=%d,%d,%d,%d,%d "
                "p=%d-%d",
                arch.c_str(), b->ncu, build.c_str(), b->stats_mode == FBH_STATS_FX64 ? "fx64" : "ref16", b->probit,
                b->W, b->CL, b->m, b->node_cap, q(0), q(1), q(2), q(3), q(4), ps.front(), ps.back());
  std::string key(buf);
  for (char& ch : key)
    if (ch == '\t' || ch == '\n' || ch == '\r') ch = '_';
  return key;
}

struct TuneEntry {
  int packing, tile_rows, resident, chains_per_cu;
  double ms;
};

// the last entry for key in the cache file; false if the file or the key is absent (malformed lines are skipped)
static bool tune_lookup(const char* path, const std::string& key, TuneEntry* e) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  bool found = false;
  char* line = nullptr;
  size_t cap = 0;
  ssize_t len;
  while ((len = getline(&line, &cap, f)) >= 0) {
    const char* tab = std::strchr(line, '\t');
    if (!tab || (size_t)(tab - line) != key.size() || std::memcmp(line, key.data(), key.size()) != 0) continue;
    TuneEntry t{};
    if (std::sscanf(tab + 1, "%d %d %d %d %lf", &t.packing, &t.tile_rows, &t.resident, &t.chains_per_cu, &t.ms) ==
        5) {
      *e = t;
      found = true;
    }
  }
  std::free(line);
  std::fclose(f);
  return found;
}

extern "C" int fbh_batch_autotune_cached(fbh_batch* b, int trial_sweeps, const char* cache_path, fbh_tune* out) {
  if (!b || trial_sweeps < 1) {
    set_error(FBH_EINVAL, "fbh_batch_autotune_cached: bad argument");
    return FBH_EINVAL;
  }
  if (!cache_path || !*cache_path) return fbh_batch_autotune(b, trial_sweeps, out);
  if (b->net) {
    set_error(FBH_EINVAL, "fbh_batch_autotune_cached: not for a batch set up for the learned proposal");
    return FBH_EINVAL;
  }
  const std::string key = tune_key(b);
  TuneEntry e{};
  if (tune_lookup(cache_path, key, &e)) {
    // an entry this batch cannot run (edited by hand, or from a build with other limits) is tuned afresh
    const bool dense = e.chains_per_cu == 2;
    const int kind = (e.chains_per_cu == 1 || e.chains_per_cu == 2) && (e.resident == 1 || e.resident == 2) &&
                             (e.tile_rows == 1024 || e.tile_rows == 1536)
                         ? pick_kind(b, e.tile_rows, e.resident, dense)
                         : -1;
    if (kind >= 0 && e.packing >= 0 && e.packing <= 2 && (b->W <= 1 || b->fits[dense][kind])) {
      const int rc = configure(b, e.packing, kind, dense);
      if (rc) return rc;
      if (out) {
        std::memset(out, 0, sizeof *out);
        out->packing = e.packing;
        out->tile_rows = 256 * kind_rows(kind);
        out->resident = kind_resident(kind) ? 2 : 1;
        out->chains_per_cu = dense ? 2 : 1;
        out->from_cache = 1;
      }
      return FBH_OK;
    }
  }
  // the file must be writable before any sweep is spent on tuning
  FILE* f = std::fopen(cache_path, "a");
  if (!f) {
    set_error(FBH_EINVAL, std::string("fbh_batch_autotune_cached: cannot open the cache file ") + cache_path);
    return FBH_EINVAL;
  }
  fbh_tune t{};
  int rc = fbh_batch_autotune(b, trial_sweeps, &t);
  if (rc) {
    std::fclose(f);
    return rc;
  }
  // the chosen configuration's time: its candidate in the packing stage (the last candidates)
  double ms = 0;
  for (int i = t.ncand - 1; i >= 0; i--)
    if (t.cand_packing[i] == t.packing && t.cand_tile_rows[i] == t.tile_rows && t.cand_resident[i] == t.resident &&
        t.cand_chains_per_cu[i] == t.chains_per_cu) {
      ms = t.ms_per_sweep[i];
      break;
    }
  char vals[128];
  std::snprintf(vals, sizeof vals, "\t%d %d %d %d %.6f\n", t.packing, t.tile_rows, t.resident, t.chains_per_cu, ms);
  const std::string line = key + vals;
  // one write of the whole line: appends from other processes do not interleave within it
  const bool ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
  if (std::fclose(f) != 0 || !ok) {
    set_error(FBH_EINVAL, std::string("fbh_batch_autotune_cached: writing the cache file ") + cache_path +
                              " failed (the tuned configuration is in effect)");
    return FBH_EINVAL;
  }
  if (out) *out = t;
  return FBH_OK;
}


// a row split whose exchange timed out (members not resident together) has invalid results: report it
static int check_split(fbh_batch* b) {
  if (b->W <= 1) return FBH_OK;
  std::vector<ChainState> cs(b->C);
  FBH_HIP_CHECK(hipMemcpy(cs.data(), b->d_state, sizeof(ChainState) * b->C, hipMemcpyDeviceToHost));
  for (int dc = 0; dc < b->C; dc++)
    if (cs[dc].xerr) {
      set_error(FBH_EHIP, "fbh_batch: a row-split exchange timed out (chain " + std::to_string(dc / b->W) +
                              "); the workgroups of a chain were not resident together");
      return FBH_EHIP;
    }
  return FBH_OK;
}

extern "C" int fbh_batch_sync(fbh_batch* b) {
  if (!b) return FBH_EINVAL;
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  // every graph launched so far has completed: release them
  for (auto g : b->gexecs) (void)hipGraphExecDestroy(g);
  for (auto g : b->graphs) (void)hipGraphDestroy(g);
  b->gexecs.clear();
  b->graphs.clear();
  return check_split(b);
}

extern "C" int fbh_batch_iter(const fbh_batch* b) { return b ? b->it : -1; }

extern "C" int fbh_batch_row_split(const fbh_batch* b) { return b ? b->W : -1; }

extern "C" int fbh_batch_fetch(fbh_batch* b, int c, const fbh_chain_out* out) {
  if (!b || !out || c < 0 || c >= b->CL) {
    set_error(FBH_EINVAL, "fbh_batch_fetch: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  int rc = check_split(b);
  if (rc) return rc;
  const int d0 = c * b->W;  // member 0 holds the chain-level outputs (every member computes the same ones)
  const ChainDesc& d = b->hdesc[d0];
  const int total = b->nskip + b->ndpost;
  if (out->sigma) FBH_HIP_CHECK(hipMemcpy(out->sigma, b->d_sigma + d.sig_off, sizeof(float) * total, hipMemcpyDeviceToHost));
  if (out->trmean) {
    for (int k = 0; k < b->W; k++) {
      const ChainDesc& dk = b->hdesc[d0 + k];
      hipLaunchKernelGGL(k_trmean, dim3((dk.n + 255) / 256), dim3(256), 0, b->stream, b->d_TM + dk.row_off,
                         b->d_trm + dk.row_off, dk.n, b->ndpost);
      FBH_HIP_CHECK(hipGetLastError());
      FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
      FBH_HIP_CHECK(hipMemcpy(out->trmean + b->row0[d0 + k], b->d_trm + dk.row_off, sizeof(float) * dk.n,
                              hipMemcpyDeviceToHost));
    }
  }
  if (out->trees && b->keep_trees)
    FBH_HIP_CHECK(hipMemcpy(out->trees, b->d_tout + d.tout_off, sizeof(uint64_t) * b->ndpost * b->m * kNodes,
                            hipMemcpyDeviceToHost));
  if (out->stats) {
    ChainState cs;
    FBH_HIP_CHECK(hipMemcpy(&cs, b->d_state + d0, sizeof cs, hipMemcpyDeviceToHost));
    out->stats[0] = cs.n_birth;
    out->stats[1] = cs.n_death;
    out->stats[2] = cs.n_cap;
  }
  return FBH_OK;
}

static uint64_t pack_host(const TreeDev& t, int s) {
  uint32_t thb;
  std::memcpy(&thb, &t.th[s], 4);
  const uint64_t vv = ((t.used >> s) & 1ull) ? (uint64_t)t.v[s] : 255u;
  return vv | ((uint64_t)t.c[s] << 8) | ((uint64_t)t.l[s] << 16) | ((uint64_t)t.r[s] << 24) | ((uint64_t)thb << 32);
}

extern "C" int fbh_batch_state(fbh_batch* b, int c, uint32_t* rng, double* sigma, double* af, double* latent,
                               uint8_t* leaf, uint64_t* words) {
  if (!b || c < 0 || c >= b->CL) {
    set_error(FBH_EINVAL, "fbh_batch_state: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  const int d0 = c * b->W;
  const ChainDesc& d = b->hdesc[d0];
  const int n_all = d.n_total;
  ChainState cs;
  FBH_HIP_CHECK(hipMemcpy(&cs, b->d_state + d0, sizeof cs, hipMemcpyDeviceToHost));
  if (rng) std::memcpy(rng, cs.rng, 16);
  if (sigma) *sigma = cs.sigma;
  for (int k = 0; k < b->W; k++) {
    const ChainDesc& dk = b->hdesc[d0 + k];
    const int r0 = b->row0[d0 + k];
    if (af) FBH_HIP_CHECK(hipMemcpy(af + r0, b->d_AF + dk.row_off, sizeof(double) * dk.n, hipMemcpyDeviceToHost));
    if (latent)
      FBH_HIP_CHECK(hipMemcpy(latent + r0, b->d_Y + dk.row_off, sizeof(double) * dk.n, hipMemcpyDeviceToHost));
    if (leaf)
      for (int j = 0; j < b->m; j++)
        FBH_HIP_CHECK(hipMemcpy(leaf + (size_t)j * n_all + r0, b->d_leaf + dk.leaf_off + (size_t)j * dk.ns, dk.n,
                                hipMemcpyDeviceToHost));
  }
  if (words) {
    std::vector<TreeDev> tr(b->m);
    FBH_HIP_CHECK(hipMemcpy(tr.data(), b->d_trees + d.tree_off, sizeof(TreeDev) * b->m, hipMemcpyDeviceToHost));
    for (int j = 0; j < b->m; j++)
      for (int s = 0; s < kNodes; s++) words[(size_t)j * kNodes + s] = pack_host(tr[j], s);
  }
  return FBH_OK;
}

extern "C" int fbh_batch_timing(const fbh_batch* b, double* avg_ms, int* nlaunch, double* bytes_per_launch) {
  if (!b) return FBH_EINVAL;
  double tot = 0;
  int k = 0;
  for (const auto& t : b->ev) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.e0, t.e1) == hipSuccess) {
      tot += ms;
      k += t.nl;
    } else {
      (void)hipGetLastError();  // a pair not yet complete: skipped, and its error not left for the next call
    }
  }
  if (avg_ms) *avg_ms = k ? tot / k : 0;
  if (nlaunch) *nlaunch = k;
  if (bytes_per_launch) *bytes_per_launch = b->bytes_per_sweep * b->sweeps_per_launch;
  return FBH_OK;
}

// debugging hook: per-tree-update records of the next launches (cap records per chain)
extern "C" int fbh_batch_trace(fbh_batch* b, int cap, void* host_out, int chain) {
  if (!b) return FBH_EINVAL;
  FBH_HIP_CHECK(hipSetDevice(b->device));
  if (cap > 0 && !b->d_trace) {
    if (dalloc(&b->d_trace, (size_t)b->C * cap)) return FBH_ENOMEM;
    FBH_HIP_CHECK(hipMemset(b->d_trace, 0, sizeof(TraceRec) * (size_t)b->C * cap));
    b->trace_cap = cap;
    return FBH_OK;
  }
  if (host_out && b->d_trace) {
    if (chain < 0 || chain >= b->CL) return FBH_EINVAL;
    FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
    FBH_HIP_CHECK(hipMemcpy(host_out, b->d_trace + (size_t)chain * b->W * b->trace_cap,
                            sizeof(TraceRec) * b->trace_cap, hipMemcpyDeviceToHost));
  }
  return FBH_OK;
}

extern "C" void fbh_batch_destroy(fbh_batch* b) { free_batch(b); }

// the device yhat walk (fbh_predict.hip)
namespace fbh {
int yhat_device(const uint64_t* d_trees, const uint8_t* d_q, int n, int ldq, int ndpost, int m, hipStream_t stream,
                double* host_out);
}

extern "C" int fbh_batch_yhat(fbh_batch* b, int c, double* yhat) {
  if (!b || c < 0 || c >= b->CL || !yhat || !b->keep_trees) {
    set_error(FBH_EINVAL, "fbh_batch_yhat: bad argument or trees not kept");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  const int d0 = c * b->W;
  const ChainDesc& d = b->hdesc[d0];
  if (b->W == 1)
    return fbh::yhat_device(b->d_tout + d.tout_off, b->d_q + d.q_off, d.n, d.ns, b->ndpost, b->m, b->stream, yhat);
  // a row split: each member's rows against the chain's kept trees, placed at the member's row offset
  const int n_all = d.n_total;
  std::vector<double> part;
  for (int k = 0; k < b->W; k++) {
    const ChainDesc& dk = b->hdesc[d0 + k];
    part.resize((size_t)b->ndpost * dk.n);
    const int rc = fbh::yhat_device(b->d_tout + d.tout_off, b->d_q + dk.q_off, dk.n, dk.ns, b->ndpost, b->m,
                                    b->stream, part.data());
    if (rc) return rc;
    for (int dr = 0; dr < b->ndpost; dr++)
      std::memcpy(yhat + (size_t)dr * n_all + b->row0[d0 + k], part.data() + (size_t)dr * dk.n, sizeof(double) * dk.n);
  }
  return FBH_OK;
}

extern "C" int fbh_fit_batch(const fbh_chain_in* chains, int nchains, const fbh_opts* opts, const fbh_chain_out* outs) {
  fbh_batch* b = nullptr;
  int rc = fbh_batch_create(chains, nchains, opts, &b);
  if (rc) return rc;
  rc = fbh_batch_run(b, opts->nskip + opts->ndpost);
  if (!rc) rc = fbh_batch_sync(b);
  for (int c = 0; !rc && c < nchains; c++) rc = fbh_batch_fetch(b, c, &outs[c]);
  fbh_batch_destroy(b);
  return rc;
}

// ==================================================================== drop-in single chain (src/bart_kernel.h:49-70)
static bool g_external_rng = false;

extern "C" void bart_set_rng(double (*u)(void), double (*nrm)(void), double (*e)(void), double (*chisq)(double)) {
  g_external_rng = u || nrm || e || chisq;
}

extern "C" void bart_fit_cpu(const uint8_t* qin, const float* yin, const float* wtin, const int* numcut,
                             const float* pgtab, float* sigma_out, float* trmean_out, uint64_t* trees_out, int n,
                             int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
                             unsigned seed, int keep_trees, int* stats_out, int probit, const int* ybin, float offset) {
  if (g_external_rng) {
    set_error(FBH_ERNG, "bart_fit_cpu: an external generator is installed (FASTBART_RNG=R); the device path only "
                        "reproduces the internal xoshiro128** stream");
    if (stats_out) stats_out[0] = stats_out[1] = stats_out[2] = -1;
    return;
  }
  fbh_chain_in ci;
  std::memset(&ci, 0, sizeof ci);
  ci.n = n;
  ci.p = p;
  ci.q = qin;
  ci.y = yin;
  ci.wt = wtin;
  ci.numcut = numcut;
  ci.ybin = ybin;
  ci.tau = tau;
  ci.lambda = lambda;
  ci.nu = nu;
  ci.sigest = sigest;
  ci.offset = offset;
  ci.seed = seed;
  fbh_opts o;
  std::memset(&o, 0, sizeof o);
  o.m = m;
  o.nskip = nskip;
  o.ndpost = ndpost;
  o.probit = probit;
  o.keep_trees = keep_trees && trees_out;
  o.pgtab = pgtab;
  fbh_chain_out co{sigma_out, trmean_out, o.keep_trees ? trees_out : nullptr, nullptr};
  int st[3] = {0, 0, 0};
  co.stats = st;
  const int rc = fbh_fit_batch(&ci, 1, &o, &co);
  if (rc) {
    if (stats_out) stats_out[0] = stats_out[1] = stats_out[2] = -1;
    return;
  }
  set_error(FBH_OK, "");
  if (stats_out) {
    stats_out[0] = st[0];
    stats_out[1] = st[1];
    stats_out[2] = st[2];
  }
}

extern "C" void bart_wfit(const uint8_t* qin, const float* yin, const float* wtin, const int* numcut,
                          const float* pgtab, float* sigma_out, float* trmean_out, uint64_t* trees_out, int n, int p,
                          int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest, unsigned seed,
                          int keep_trees, int* stats_out) {
  bart_fit_cpu(qin, yin, wtin, numcut, pgtab, sigma_out, trmean_out, trees_out, n, p, m, nskip, ndpost, tau, lambda,
               nu, sigest, seed, keep_trees, stats_out, 0, nullptr, 0.f);
}

// ==================================================================== learned birth proposal (host)
static LpArgs lp_args(const fbh_batch* b, int j) {
  LpArgs A;
  A.a = make_args(b);
  A.img = b->d_img;
  A.st = b->d_lps;
  A.AFX = b->d_afx;
  A.X = b->d_X;
  A.rg = b->d_rg;
  int ld = 0;
  A.Z = b->net ? mlp_logits(b->net, &ld) : nullptr;
  A.ldz = ld;
  A.alpha = b->alpha;
  A.px = b->d_px;
  A.prg = b->d_prg;
  A.ptg = b->d_ptg;
  A.pcount = b->d_pcount;
  A.pcap = b->pcap;
  A.pent = b->d_pent;
  A.pq = b->d_pq;
  A.tseed = b->tseed;
  A.keep = b->keep;
  A.target = b->target;
  A.j = j;
  return A;
}

extern "C" int fbh_batch_learned_setup(fbh_batch* b, fbh_mlp* net, float alpha, long pair_capacity) {
  if (!b || !(alpha >= 0.f && alpha <= 1.f) || (alpha > 0.f && !net) || pair_capacity < 0 ||
      (pair_capacity > 0 && !net)) {
    set_error(FBH_EINVAL, "fbh_batch_learned_setup: bad argument (alpha in [0, 1], a network when alpha > 0 or when "
                          "collecting pairs)");
    return FBH_EINVAL;
  }
  if (b->probit || b->W > 1) {
    set_error(FBH_EINVAL, "fbh_batch_learned_setup: the learned proposal runs wbart chains without a row split only");
    return FBH_EINVAL;
  }
  if (net) {
    int d[4], mb = 0;
    mlp_dims(net, d, &mb);
    if (d[0] != kFeat || d[3] != kClasses || mb < b->C) {
      set_error(FBH_EINVAL, "fbh_batch_learned_setup: the network must map 513 features to 8160 classes with a batch "
                            "of at least the chain count");
      return FBH_EINVAL;
    }
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  int rc;
  if (!b->d_img) {
    if ((rc = dalloc(&b->d_img, b->C)) || (rc = dalloc(&b->d_lps, b->C)) || (rc = dalloc(&b->d_afx, b->rows_total)) ||
        (rc = dalloc(&b->d_X, (size_t)b->C * kFeat)) || (rc = dalloc(&b->d_rg, (size_t)b->C * FBH_PMAX)) ||
        (rc = dalloc(&b->d_pcount, 1)))
      return rc;
    FBH_HIP_CHECK(hipMemset(b->d_pcount, 0, sizeof(unsigned long long)));
    FBH_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_lp_stage1<false>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, Geo<kRowsStreamL>::kScratchBytes));
  }
  if (pair_capacity != b->pcap) {
    for (void* q : {(void*)b->d_px, (void*)b->d_prg, (void*)b->d_ptg, (void*)b->d_idx, (void*)b->d_pent,
                    (void*)b->d_pq})
      if (q) (void)hipFree(q);
    b->d_pq = nullptr;
    b->hold = 0;
    b->d_px = nullptr;
    b->d_prg = nullptr;
    b->d_ptg = nullptr;
    b->d_idx = nullptr;
    b->d_pent = nullptr;
    b->pcap = 0;
    if (pair_capacity > 0) {
      int d[4], mb = 0;
      mlp_dims(net, d, &mb);
      if ((rc = dalloc(&b->d_px, (size_t)pair_capacity * kFeat)) ||
          (rc = dalloc(&b->d_prg, (size_t)pair_capacity * FBH_PMAX)) || (rc = dalloc(&b->d_ptg, pair_capacity)) ||
          (rc = dalloc(&b->d_idx, mb)) || (rc = dalloc(&b->d_pent, (size_t)pair_capacity * 2)))
        return rc;
      b->pcap = pair_capacity;
      if (b->target == 1 && (rc = dalloc(&b->d_pq, (size_t)pair_capacity * kClasses))) return rc;
    }
    FBH_HIP_CHECK(hipMemset(b->d_pcount, 0, sizeof(unsigned long long)));
  }
  b->net = net;
  b->alpha = alpha;
  return FBH_OK;
}

template <bool FX>
static int lp_run(fbh_batch* b, int nsweeps) {
  const size_t dyn = Geo<kRowsStreamL>::kScratchBytes;
  for (auto fn : {reinterpret_cast<const void*>(k_lp_stage1<FX>), reinterpret_cast<const void*>(k_lp_stage2<FX>),
                  reinterpret_cast<const void*>(k_lp_close<FX>)})
    FBH_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
  hipLaunchKernelGGL(k_lp_begin, dim3(b->C), dim3(kBlock), 0, b->stream, lp_args(b, 0));
  FBH_HIP_CHECK(hipGetLastError());
  for (int sw = 0; sw < nsweeps; sw++) {
    hipEvent_t e0, e1;
    FBH_HIP_CHECK(hipEventCreate(&e0));
    FBH_HIP_CHECK(hipEventCreate(&e1));
    b->lev.push_back(e0);
    b->lev.push_back(e1);
    FBH_HIP_CHECK(hipEventRecord(e0, b->stream));
    for (int j = 0; j < b->m; j++) {
      const LpArgs A = lp_args(b, j);
      hipLaunchKernelGGL(k_lp_stage1<FX>, dim3(b->C), dim3(kBlock), dyn, b->stream, A);
      if (b->target == 1 && b->pcap > 0)
        hipLaunchKernelGGL(k_lp_target, dim3(b->C), dim3(kTgtBlock), 0, b->stream, A);
      if (b->alpha > 0.f) {
        const int rc = mlp_forward_dev(b->net, b->d_X, b->C, b->stream);
        if (rc) return rc;
      }
      hipLaunchKernelGGL(k_lp_stage2<FX>, dim3(b->C), dim3(kBlock), dyn, b->stream, A);
    }
    hipLaunchKernelGGL(k_lp_close<FX>, dim3(b->C), dim3(kBlock), dyn, b->stream, lp_args(b, b->m - 1));
    FBH_HIP_CHECK(hipGetLastError());
    FBH_HIP_CHECK(hipEventRecord(e1, b->stream));
  }
  hipLaunchKernelGGL(k_lp_end, dim3((b->C + 63) / 64), dim3(64), 0, b->stream, lp_args(b, 0), b->C);
  FBH_HIP_CHECK(hipGetLastError());
  b->it += nsweeps;
  return FBH_OK;
}

extern "C" int fbh_batch_run_learned(fbh_batch* b, int nsweeps) {
  if (!b || nsweeps < 0) {
    set_error(FBH_EINVAL, "fbh_batch_run_learned: bad argument");
    return FBH_EINVAL;
  }
  if (!b->d_img) {
    set_error(FBH_EINVAL, "fbh_batch_run_learned: call fbh_batch_learned_setup first");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  nsweeps = std::min(nsweeps, b->nskip + b->ndpost - b->it);
  if (nsweeps <= 0) return FBH_OK;
  return b->stats_mode == FBH_STATS_FX64 ? lp_run<true>(b, nsweeps) : lp_run<false>(b, nsweeps);
}

extern "C" int fbh_batch_learned_target(fbh_batch* b, int kind, float fraction, unsigned long long seed) {
  if (!b || (kind != 0 && kind != 1) || !(fraction > 0.f && fraction <= 1.f)) {
    set_error(FBH_EINVAL, "fbh_batch_learned_target: kind must be 0 (accepted births, BART's rule) or 1 (proposed "
                          "births, the exact split conditional), fraction in (0, 1]");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  b->target = kind;
  b->keep = fraction;
  b->tseed = seed;
  if (kind == 1 && b->pcap > 0 && !b->d_pq) {
    const int rc = dalloc(&b->d_pq, (size_t)b->pcap * kClasses);
    if (rc) return rc;
  }
  return FBH_OK;
}

extern "C" int fbh_batch_learned_holdout(fbh_batch* b, long count) {
  if (!b || count < 0) {
    set_error(FBH_EINVAL, "fbh_batch_learned_holdout: bad argument");
    return FBH_EINVAL;
  }
  b->hold = count;
  return FBH_OK;
}

extern "C" int fbh_batch_learned_eval(fbh_batch* b, long k0, long count, double* mean_loss, double* mean_entropy,
                                      double* mean_log_allowed) {
  const long np = fbh_batch_learned_pairs(b);
  if (!b || !b->net || np < 0 || k0 < 0 || count < 1 || k0 + count > np || !mean_loss) {
    set_error(FBH_EINVAL, "fbh_batch_learned_eval: no network or a range outside the collected pairs");
    return FBH_EINVAL;
  }
  int d[4], mb = 0;
  mlp_dims(b->net, d, &mb);
  float *x, *loss;
  int* tgt;
  uint8_t* mask;
  mlp_batch_buffers(b->net, &x, &tgt, &mask, &loss);
  const bool soft = b->target == 1 && b->d_pq;
  float* sb = soft ? mlp_soft_buffer(b->net) : nullptr;
  if (soft && !sb) return FBH_ENOMEM;
  std::vector<int> idx(mb);
  std::vector<float> l(mb);
  double tot = 0;
  for (long k = k0; k < k0 + count; k += mb) {
    const int M = (int)std::min<long>(mb, k0 + count - k);
    for (int t = 0; t < M; t++) idx[t] = (int)(k + t);
    FBH_HIP_CHECK(hipMemcpyAsync(b->d_idx, idx.data(), sizeof(int) * M, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_lp_gather, dim3(M), dim3(256), 0, b->stream, b->d_px, b->d_prg, b->d_ptg,
                       soft ? b->d_pq : nullptr, b->d_idx, M, x, tgt, mask, sb);
    FBH_HIP_CHECK(hipGetLastError());
    int rc;
    if ((rc = mlp_forward_dev(b->net, x, M, b->stream)) || (rc = mlp_loss_dev(b->net, M, soft, b->stream))) return rc;
    FBH_HIP_CHECK(hipMemcpyAsync(l.data(), loss, sizeof(float) * M, hipMemcpyDeviceToHost, b->stream));
    FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
    for (int t = 0; t < M; t++) tot += l[t];
  }
  *mean_loss = tot / count;
  if (mean_entropy || mean_log_allowed) {
    double e = std::nan(""), la = std::nan("");
    if (b->target == 1) {
      std::vector<float> h((size_t)count * 2);
      FBH_HIP_CHECK(hipMemcpy(h.data(), b->d_pent + 2 * k0, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
      e = la = 0;
      for (long k = 0; k < count; k++) {
        e += h[2 * k];
        la += h[2 * k + 1];
      }
      e /= count;
      la /= count;
    }
    if (mean_entropy) *mean_entropy = e;
    if (mean_log_allowed) *mean_log_allowed = la;
  }
  return FBH_OK;
}

extern "C" int fbh_batch_learned_target_stats(fbh_batch* b, double* mean_entropy, double* mean_log_allowed) {
  if (!b || !mean_entropy || !mean_log_allowed) {
    set_error(FBH_EINVAL, "fbh_batch_learned_target_stats: bad argument");
    return FBH_EINVAL;
  }
  *mean_entropy = *mean_log_allowed = std::nan("");
  const long np = fbh_batch_learned_pairs(b);
  if (np < 0) {
    set_error(FBH_EHIP, "fbh_batch_learned_target_stats: reading the pair count failed");
    return FBH_EHIP;
  }
  if (b->target != 1 || np == 0) return FBH_OK;
  std::vector<float> h((size_t)np * 2);
  FBH_HIP_CHECK(hipMemcpy(h.data(), b->d_pent, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
  double e = 0, l = 0;
  for (long k = 0; k < np; k++) {
    e += h[2 * k];
    l += h[2 * k + 1];
  }
  *mean_entropy = e / np;
  *mean_log_allowed = l / np;
  return FBH_OK;
}

extern "C" int fbh_batch_learned_pairs_copy(fbh_batch* b, long k0, long count, float* features, uint16_t* ranges,
                                            int* cls, float* entropy) {
  const long np = fbh_batch_learned_pairs(b);
  if (!b || np < 0 || k0 < 0 || count < 0 || k0 + count > np) {
    set_error(FBH_EINVAL, "fbh_batch_learned_pairs_copy: range outside the collected pairs");
    return FBH_EINVAL;
  }
  if (count == 0) return FBH_OK;
  if (features)
    FBH_HIP_CHECK(hipMemcpy(features, b->d_px + (size_t)k0 * kFeat, sizeof(float) * kFeat * count,
                            hipMemcpyDeviceToHost));
  if (ranges)
    FBH_HIP_CHECK(hipMemcpy(ranges, b->d_prg + (size_t)k0 * FBH_PMAX, sizeof(uint16_t) * FBH_PMAX * count,
                            hipMemcpyDeviceToHost));
  if (cls) FBH_HIP_CHECK(hipMemcpy(cls, b->d_ptg + k0, sizeof(int) * count, hipMemcpyDeviceToHost));
  if (entropy) {
    if (b->target == 1) {
      FBH_HIP_CHECK(hipMemcpy(entropy, b->d_pent + 2 * k0, sizeof(float) * 2 * count, hipMemcpyDeviceToHost));
    } else {
      for (long k = 0; k < 2 * count; k++) entropy[k] = std::nanf("");
    }
  }
  return FBH_OK;
}

extern "C" int fbh_debug_split_conditional(int n, int p, const uint8_t* q, const uint8_t* member, const double* w,
                                           const double* r, const uint16_t* ranges, int depth, int ngv, int ngood,
                                           const double* pg, double sigma, double tau, const double* u, int nu,
                                           float* q_out, int* cls_out, double* entropy) {
  if (n < 1 || p < 1 || p > FBH_PMAX || !q || !member || !w || !r || !ranges || depth < 0 || depth + 1 >= kNodes ||
      !pg || !(sigma > 0) || !(tau > 0) || nu < 0 || (nu > 0 && (!u || !cls_out)) || !q_out || !entropy) {
    set_error(FBH_EINVAL, "fbh_debug_split_conditional: bad argument");
    return FBH_EINVAL;
  }
  for (int v = 0; v < p; v++) {
    const int L = ranges[v] & 255, U = ranges[v] >> 8;
    if (U >= L && U > 254) {
      set_error(FBH_EINVAL, "fbh_debug_split_conditional: cut above 254");
      return FBH_EINVAL;
    }
  }
  int rc = ensure_device(0);
  if (rc) return rc;
  uint8_t *dq = nullptr, *dm = nullptr;
  double *dw = nullptr, *dr = nullptr, *du = nullptr, *dpg = nullptr, *dh = nullptr;
  float* dqo = nullptr;
  int* dc = nullptr;
  auto cleanup = [&]() {
    for (void* x : {(void*)dq, (void*)dm, (void*)dw, (void*)dr, (void*)du, (void*)dpg, (void*)dh, (void*)dqo,
                    (void*)dc})
      if (x) (void)hipFree(x);
  };
  if ((rc = dalloc(&dq, (size_t)n * p)) || (rc = dalloc(&dm, n)) || (rc = dalloc(&dw, n)) || (rc = dalloc(&dr, n)) ||
      (rc = dalloc(&du, std::max(nu, 1))) || (rc = dalloc(&dpg, kNodes)) || (rc = dalloc(&dh, 1)) ||
      (rc = dalloc(&dqo, (size_t)p * 256)) || (rc = dalloc(&dc, std::max(nu, 1)))) {
    cleanup();
    return rc;
  }
  hipError_t e = hipSuccess;
  auto up = [&](void* d, const void* h, size_t bytes) {
    if (e == hipSuccess && bytes) e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
  };
  up(dq, q, (size_t)n * p);
  up(dm, member, n);
  up(dw, w, sizeof(double) * n);
  up(dr, r, sizeof(double) * n);
  up(dpg, pg, sizeof(double) * kNodes);
  if (nu) up(du, u, sizeof(double) * nu);
  SplitNode N{};
  N.n = n;
  N.p = p;
  N.ns = n;
  N.d = depth;
  N.ngv = ngv;
  N.ngood = ngood;
  N.inv_s2 = 1.0 / (sigma * sigma);
  N.t2 = tau * tau;
  N.PG = dpg;
  N.Q = dq;
  for (int v = 0; v < FBH_PMAX; v++) N.rg[v] = v < p ? ranges[v] : kNoRange;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_debug_split, dim3(1), dim3(kTgtBlock), 0, nullptr, N, dm, dw, dr, du, nu, dqo, dc, dh);
    e = hipGetLastError();
  }
  std::vector<float> qq((size_t)p * 256);
  if (e == hipSuccess) e = hipMemcpy(qq.data(), dqo, sizeof(float) * qq.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(entropy, dh, sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && nu) e = hipMemcpy(cls_out, dc, sizeof(int) * nu, hipMemcpyDeviceToHost);
  cleanup();
  if (e != hipSuccess) {
    set_error(FBH_EHIP, std::string("fbh_debug_split_conditional: ") + hipGetErrorString(e));
    return FBH_EHIP;
  }
  for (int v = 0; v < p; v++)
    for (int k = 0; k < kCuts; k++) q_out[(size_t)v * kCuts + k] = qq[(size_t)v * 256 + k];
  return FBH_OK;
}

extern "C" long fbh_batch_learned_pairs(fbh_batch* b) {
  if (!b || !b->d_pcount) return 0;
  if (hipSetDevice(b->device) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) return -1;
  unsigned long long k = 0;
  if (hipMemcpy(&k, b->d_pcount, sizeof k, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  return (long)std::min<unsigned long long>(k, (unsigned long long)b->pcap);
}

extern "C" int fbh_batch_learned_train(fbh_batch* b, int steps, int batch, float lr, unsigned seed, float* loss_out) {
  if (!b || !b->net || steps < 0 || batch < 1 || !(lr > 0.f)) {
    set_error(FBH_EINVAL, "fbh_batch_learned_train: bad argument or no network");
    return FBH_EINVAL;
  }
  int d[4], mb = 0;
  mlp_dims(b->net, d, &mb);
  if (batch > mb) {
    set_error(FBH_EINVAL, "fbh_batch_learned_train: batch above the network's max_batch");
    return FBH_EINVAL;
  }
  const long npairs = fbh_batch_learned_pairs(b) - b->hold;
  if (npairs < 1) {
    set_error(FBH_EINVAL, "fbh_batch_learned_train: no training pairs collected (past the held-out ones)");
    return FBH_EINVAL;
  }
  float *x, *loss;
  int* tgt;
  uint8_t* mask;
  mlp_batch_buffers(b->net, &x, &tgt, &mask, &loss);
  const bool soft = b->target == 1 && b->d_pq;
  float* sb = soft ? mlp_soft_buffer(b->net) : nullptr;
  if (soft && !sb) return FBH_ENOMEM;
  uint64_t st = (uint64_t)seed * 0x9E3779B97F4A7C15ull + 3;
  std::vector<int> idx(batch);
  std::vector<float> l(batch);
  for (int s = 0; s < steps; s++) {
    for (int k = 0; k < batch; k++) {
      uint64_t z = (st += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      idx[k] = (int)((z ^ (z >> 31)) % (uint64_t)npairs);
    }
    FBH_HIP_CHECK(hipMemcpyAsync(b->d_idx, idx.data(), sizeof(int) * batch, hipMemcpyHostToDevice, b->stream));
    hipLaunchKernelGGL(k_lp_gather, dim3(batch), dim3(256), 0, b->stream, b->d_px, b->d_prg, b->d_ptg,
                       soft ? b->d_pq : nullptr, b->d_idx, batch, x, tgt, mask, sb);
    FBH_HIP_CHECK(hipGetLastError());
    int rc;
    if ((rc = mlp_forward_dev(b->net, x, batch, b->stream)) ||
        (rc = mlp_backward_dev(b->net, batch, b->stream, soft)) ||
        (rc = mlp_adam_dev(b->net, lr, b->stream)))
      return rc;
    if (loss_out) {
      FBH_HIP_CHECK(hipMemcpyAsync(l.data(), loss, sizeof(float) * batch, hipMemcpyDeviceToHost, b->stream));
      FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
      double t = 0;
      for (float v : l) t += v;
      loss_out[s] = (float)(t / batch);
    }
  }
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  return FBH_OK;
}

extern "C" int fbh_batch_learned_timing(const fbh_batch* b, double* avg_ms, int* nsweeps) {
  if (!b) return FBH_EINVAL;
  double tot = 0;
  int k = 0;
  for (size_t i = 0; i + 1 < b->lev.size(); i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, b->lev[i], b->lev[i + 1]) == hipSuccess) {
      tot += ms;
      k++;
    }
  }
  if (avg_ms) *avg_ms = k ? tot / k : 0;
  if (nsweeps) *nsweeps = k;
  return FBH_OK;
}
