// The batched BART sampler on gfx950: one persistent workgroup per chain runs whole sweeps on the device.
//
// Per tree update (src/bart_kernel.cpp:400-535): proposal, one data pass over the chain's rows (apply the previous
// tree's accepted change to the leaf ids and the running fit, then this tree's sufficient statistics), the
// Metropolis-Hastings decision, and the leaf draws. After the last tree, a closing pass (apply, residual sum of squares,
// posterior-mean sum) and the sigma draw or the probit latent redraw (:536-558).
//
// Data pass (the hot loop, :133-213): rows are staged through LDS in tiles. Phase A is row-parallel over all threads
// (coalesced reads of the u8 leaf-id and cutpoint-rank columns, the running fit, y and w; leaf-id rewrite) and writes
// per-row w, w*r and a tag (leaf rank | birth-split side). Phase B reduces per (lane, leaf): in REF16 mode each lane
// i % 16 is summed in row order by one thread with a branch-free fused multiply-add by 1.0 or 0.0 (adding +0.0 to a sum
// that started at +0.0 leaves it unchanged, so the order and rounding are exactly the reference's lane sums), then the
// 16 lanes are added 0..15. In FX64 mode per-row int64 fixed-point values are added with LDS atomics (order free).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

constexpr int kThreads = 256;
constexpr int kTile = 2048;  // rows per LDS tile (multiple of 16)
constexpr int kGroups = (kMaxRank + kThreads / kLanes - 1) / (kThreads / kLanes);  // rank groups per phase-B thread

struct KArgs {
  const ChainDesc* desc;
  ChainState* state;
  TreeDev* trees;
  const uint8_t* q;
  double* Y;
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
};

struct Rng {
  uint32_t s0, s1, s2, s3;
};

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
// (k + 1/2) 2^-24 from the top 24 bits (:59)
__device__ __forceinline__ float unif_f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
// sqrtf(-2 logf(u1)) cosf(2 pi u2) in float with the host libm images (:60)
__device__ __forceinline__ float normal_f(uint32_t x1, uint32_t x2, const LibmTables& tb) {
  return __fsqrt_rn(-2.f * tb.logf_t[x1 >> 8]) * tb.cosf_t[x2 >> 8];
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
  const double c = 1.0 / __dsqrt_rn(9.0 * d);
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
    const double lam = 0.5 * (tau + __dsqrt_rn(tau * tau + 4.0));
    for (;;) {
      const double e = -log_unif(g, tb);
      z = e / lam + tau;
      const double u = unif_d(g);
      if (!(u > exp(-0.5 * (z - lam) * (z - lam)))) break;
    }
  }
  return z + mean;
}

// ------------------------------------------------------------------ shared memory of one chain's workgroup
struct Smem {
  TreeDev T;
  ChainDesc cd;
  double PG[kNodes];
  double thc[kNodes], tho[kNodes], thn[kNodes];
  double S0[kNodes], S1[kNodes];
  double part0[kMaxRank][kLanes], part1[kMaxRank][kLanes];
  unsigned long long fx0[kNodes], fx1[kNodes], fxs[4], fxrss;
  uint32_t rnd[132];
  uint8_t rank_of[kNodes], slot_of[kNodes];
  int pmode, pnx, pv, pc, pL, pR, pl, pr;
  int cmode, cnx, cv, cc;
  int nleaf, nL, nR;
  double prr, inv_s2, sL0, sL1, sR0, sR1, rss;
  double wsc[kTile], wrsc[kTile];
  uint8_t tag[kTile];
};

#define BITM(mask, i) (((mask) >> (i)) & 1ull)

__device__ __forceinline__ int select_kth(uint64_t m, int k) {
  int seen = 0;
  for (int i = 0; i < kNodes; i++)
    if (BITM(m, i)) {
      if (seen == k) return i;
      seen++;
    }
  return 0;
}
__device__ __forceinline__ int first_free(uint64_t used, int skip, int cap) {
  for (int i = 1; i < cap; i++)
    if (!BITM(used, i) && i != skip) return i;
  return -1;
}
// BART's tree::rg (:238-251)
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
// variables with a non-empty range at node x (:253-276)
__device__ int count_goodvars(const TreeDev& t, int x, const int* nc, int p) {
  int lo[FBH_PMAX], hi[FBH_PMAX];
  for (int v = 0; v < FBH_PMAX; v++) {
    lo[v] = 0;
    hi[v] = nc[v] - 1;
  }
  int node = x;
  for (int s = 0; s < kMaxDepth && node != 0; s++) {
    const int pa = t.par[node], v = t.v[pa], c = t.c[pa];
    if ((int)t.l[pa] == node) {
      if (c - 1 < hi[v]) hi[v] = c - 1;
    } else {
      if (c + 1 > lo[v]) lo[v] = c + 1;
    }
    node = pa;
  }
  int gcount = 0;
  for (int v = 0; v < p; v++) gcount += hi[v] >= lo[v];
  return gcount;
}
__device__ __forceinline__ double lh(double b, double M, double t2) {
  const double k = b * t2 + 1;
  return -0.5 * log(k) + 0.5 * M * M * t2 / k;
}

// Birth/death proposal for the tree in sm.T (:413-464). Runs on one thread.
__device__ void propose(Smem& sm, Rng& g, int& ngood_out, int& nnogs_out) {
  const TreeDev& t = sm.T;
  const double pb = 0.5;
  const int p = sm.cd.p;
  uint64_t goodm = 0, nogm = 0;
  for (int s = 0; s < kNodes; s++) {
    const bool isleaf = BITM(t.leaf, s), isused = BITM(t.used, s);
    if (isused && isleaf && t.ngv[s] > 0) goodm |= 1ull << s;
    if (isused && !isleaf && BITM(t.leaf, t.l[s]) && BITM(t.leaf, t.r[s])) nogm |= 1ull << s;
  }
  const int ngood = __popcll(goodm), nnogs = __popcll(nogm);
  ngood_out = ngood;
  nnogs_out = nnogs;
  const bool single = t.used == 1ull;
  const double PBx = ngood == 0 ? 0 : (single ? 1 : pb);
  int cmode = 0, cnx = 0, cv = 0, cc = 0;
  double pr = 0;
  if (unif_d(g) < PBx) {
    const int nx = select_kth(goodm, (int)(unif_d(g) * ngood));
    int v = (int)(unif_d(g) * p);
    if (v >= p) v = p - 1;
    int bL, bU;
    node_range(t, nx, v, sm.cd.nc[v], &bL, &bU);
    if (bU >= bL && (int)t.d[nx] + 1 < kMaxDepth) {
      int c = bL + (int)(unif_d(g) * (bU - bL + 1));
      if (c > bU) c = bU;
      const int dnx = t.d[nx];
      const double PGnx = sm.PG[dnx];
      double PGly, PGry;
      if (t.ngv[nx] > 1) {
        PGly = PGry = sm.PG[dnx + 1];
      } else {
        PGly = (c - 1 < bL) ? 0 : sm.PG[dnx + 1];
        PGry = (bU < c + 1) ? 0 : sm.PG[dnx + 1];
      }
      const double PDy = (ngood > 1) ? 1 - pb : ((PGry == 0 && PGly == 0) ? 1 : 1 - pb);
      double Pnogy;
      if (nx == 0) {
        Pnogy = 1;
      } else {
        const int pa = t.par[nx];
        const bool panog = BITM(t.leaf, t.l[pa]) && BITM(t.leaf, t.r[pa]);
        Pnogy = panog ? 1.0 / nnogs : 1.0 / (nnogs + 1);
      }
      const double Pbotx = 1.0 / ngood;
      pr = (PGnx * (1 - PGly) * (1 - PGry) * PDy * Pnogy) / ((1 - PGnx) * Pbotx * PBx);
      cmode = 1;
      cnx = nx;
      cv = v;
      cc = c;
    }
  } else {
    const int nx = select_kth(nogm, (int)(unif_d(g) * nnogs));
    const int dny = t.d[nx], l = t.l[nx], r = t.r[nx];
    const double PGny = sm.PG[dny];
    const double PGlx = t.ngv[l] > 0 ? sm.PG[dny + 1] : 0, PGrx = t.ngv[r] > 0 ? sm.PG[dny + 1] : 0;
    const double PBy = (nx == 0) ? 1 : pb;
    const int ng2 = ngood - (t.ngv[l] > 0) - (t.ngv[r] > 0) + 1;
    const double Pboty = 1.0 / ng2, PDx = 1 - PBx, Pnogx = 1.0 / nnogs;
    pr = ((1 - PGny) * PBy * Pboty) / (PGny * (1 - PGlx) * (1 - PGrx) * PDx * Pnogx);
    cmode = 2;
    cnx = nx;
  }
  sm.cmode = cmode;
  sm.cnx = cnx;
  sm.cv = cv;
  sm.cc = cc;
  sm.prr = pr;
}

__device__ __forceinline__ long long fx_round(double x, double scale) { return (long long)rint(x * scale); }

// One pass over the chain's rows (:133-213). All threads of the workgroup call it; it ends synchronised.
__device__ void data_pass(const KArgs& a, Smem& sm, double* AF, double* TM, const double* Y, const float* W,
                          const uint8_t* Q, uint8_t* LEAF, int jp, int j, bool apply, bool stats, bool rss,
                          bool keep) {
  const int tid = threadIdx.x;
  const int n = sm.cd.n;
  const bool fx = a.stats_mode == FBH_STATS_FX64;
  const int pmode = apply ? sm.pmode : 0, pnx = sm.pnx, pc = sm.pc, pL = sm.pL, pR = sm.pR, pl = sm.pl, pr = sm.pr;
  const int cmode = stats ? sm.cmode : 0, cnx = sm.cnx, cc = sm.cc;
  const int nleaf = sm.nleaf;
  const int ntask = rss ? 1 : nleaf + (cmode == 1 ? 2 : 0);
  uint8_t* lprev = apply ? LEAF + (size_t)jp * n : nullptr;
  const uint8_t* lcur = stats ? LEAF + (size_t)j * n : nullptr;
  const uint8_t* qp = (apply && pmode == 1) ? Q + (size_t)sm.pv * n : nullptr;
  const uint8_t* qc = (stats && cmode == 1) ? Q + (size_t)sm.cv * n : nullptr;

  if (tid < kNodes) {
    sm.S0[tid] = 0;
    sm.S1[tid] = 0;
    sm.fx0[tid] = 0;
    sm.fx1[tid] = 0;
  }
  if (tid < 4) sm.fxs[tid] = 0;
  if (tid == 0) {
    sm.nL = sm.nR = 0;
    sm.fxrss = 0;
    sm.sL0 = sm.sL1 = sm.sR0 = sm.sR1 = 0;
    sm.rss = 0;
  }
  // phase-B task of this thread: lane k, ranks qs + 16 g
  const int k = tid & (kLanes - 1), qs = tid / kLanes;
  double a0[kGroups], a1[kGroups];
  int qmask[kGroups], qval[kGroups];
  bool qon[kGroups];
#pragma unroll
  for (int g = 0; g < kGroups; g++) {
    const int q = qs + (kThreads / kLanes) * g;
    a0[g] = 0;
    a1[g] = 0;
    qon[g] = !fx && q < ntask;
    if (rss) {
      qmask[g] = 0;
      qval[g] = 0;
    } else if (q < nleaf) {
      qmask[g] = 63;
      qval[g] = q;
    } else {
      qmask[g] = 0xC0;
      qval[g] = q == nleaf ? 0x40 : 0x80;
    }
  }
  __syncthreads();

  for (int tb = 0; tb < n; tb += kTile) {
    const int rows = min(kTile, n - tb);
    int my_nl = 0, my_nr = 0;
    // ---- phase A: row-parallel apply + per-row values
    for (int rr = tid; rr < rows; rr += kThreads) {
      const int i = tb + rr;
      double af = AF[i];
      if (apply) {
        const int lf = lprev[i];
        int lf2 = lf;
        if (pmode == 1 && lf == pnx)
          lf2 = qp[i] <= pc ? pL : pR;
        else if (pmode == 2 && (lf == pl || lf == pr))
          lf2 = pnx;
        af = af - sm.tho[lf] + sm.thn[lf2];
        lprev[i] = (uint8_t)lf2;
        AF[i] = af;
      }
      if (rss) {
        const double e = Y[i] - af;
        const double c2 = ((double)W[i] * e) * e;
        if (fx) {
          atomicAdd(&sm.fxrss, (unsigned long long)fx_round(c2, 16777216.0));
        } else {
          sm.wsc[rr] = 0.0;
          sm.wrsc[rr] = c2;
          sm.tag[rr] = 0;
        }
        if (keep) TM[i] += af;
      } else if (stats) {
        const int lf = lcur[i];
        const double r = Y[i] - af + sm.thc[lf];
        const double w = (double)W[i];
        const double wr = w * r;
        int side = 0;
        if (cmode == 1 && lf == cnx) {
          side = qc[i] <= cc ? 1 : 2;
          my_nl += side == 1;
          my_nr += side == 2;
        }
        if (fx) {
          const unsigned long long iw = (unsigned long long)fx_round(w, 4294967296.0);
          const unsigned long long iwr = (unsigned long long)fx_round(wr, 4294967296.0);
          atomicAdd(&sm.fx0[lf], iw);
          atomicAdd(&sm.fx1[lf], iwr);
          if (side) {
            atomicAdd(&sm.fxs[side == 1 ? 0 : 2], iw);
            atomicAdd(&sm.fxs[side == 1 ? 1 : 3], iwr);
          }
        } else {
          sm.wsc[rr] = w;
          sm.wrsc[rr] = wr;
          sm.tag[rr] = (uint8_t)(sm.rank_of[lf] | (side << 6));
        }
      }
    }
    if (my_nl) atomicAdd(&sm.nL, my_nl);
    if (my_nr) atomicAdd(&sm.nR, my_nr);
    __syncthreads();
    // ---- phase B: lane-ordered sums
    if (!fx) {
#pragma unroll
      for (int g = 0; g < kGroups; g++) {
        if (!qon[g]) continue;
        const int mk = qmask[g], vl = qval[g];
        double x0 = a0[g], x1 = a1[g];
        for (int rr = k; rr < rows; rr += kLanes) {
          const double mm = ((int)(sm.tag[rr] & mk) == vl) ? 1.0 : 0.0;
          x0 = fma(mm, sm.wsc[rr], x0);
          x1 = fma(mm, sm.wrsc[rr], x1);
        }
        a0[g] = x0;
        a1[g] = x1;
      }
    }
    __syncthreads();
  }

  if (!fx) {
#pragma unroll
    for (int g = 0; g < kGroups; g++)
      if (qon[g]) {
        const int q = qs + (kThreads / kLanes) * g;
        sm.part0[q][k] = a0[g];
        sm.part1[q][k] = a1[g];
      }
    __syncthreads();
    if (tid < ntask) {
      const int q = tid;
      double x0 = 0, x1 = 0;
      for (int kk = 0; kk < kLanes; kk++) {
        x0 += sm.part0[q][kk];
        x1 += sm.part1[q][kk];
      }
      if (rss) {
        sm.rss = x1;
      } else if (q < nleaf) {
        const int s = sm.slot_of[q];
        sm.S0[s] = x0;
        sm.S1[s] = x1;
      } else if (q == nleaf) {
        sm.sL0 = x0;
        sm.sL1 = x1;
      } else {
        sm.sR0 = x0;
        sm.sR1 = x1;
      }
    }
  } else {
    if (tid < kNodes) {
      sm.S0[tid] = (double)(long long)sm.fx0[tid] * (1.0 / 4294967296.0);
      sm.S1[tid] = (double)(long long)sm.fx1[tid] * (1.0 / 4294967296.0);
    }
    if (tid == 0) {
      if (cmode == 1) {
        sm.sL0 = (double)(long long)sm.fxs[0] * (1.0 / 4294967296.0);
        sm.sL1 = (double)(long long)sm.fxs[1] * (1.0 / 4294967296.0);
        sm.sR0 = (double)(long long)sm.fxs[2] * (1.0 / 4294967296.0);
        sm.sR1 = (double)(long long)sm.fxs[3] * (1.0 / 4294967296.0);
      }
      if (rss) sm.rss = (double)(long long)sm.fxrss * (1.0 / 16777216.0);
    }
  }
  __syncthreads();
}

__device__ __forceinline__ uint64_t pack_word(const TreeDev& t, int s) {
  const uint32_t thb = __float_as_uint(t.th[s]);
  const uint64_t vv = BITM(t.used, s) ? (uint64_t)t.v[s] : 255u;
  return vv | ((uint64_t)t.c[s] << 8) | ((uint64_t)t.l[s] << 16) | ((uint64_t)t.r[s] << 24) | ((uint64_t)thb << 32);
}

__global__ __launch_bounds__(kThreads) void k_sweeps(KArgs a) {
  __shared__ Smem sm;
  const int tid = threadIdx.x;
  const int chain = blockIdx.x;
  if (tid == 0) sm.cd = a.desc[chain];
  if (tid < kNodes) sm.PG[tid] = a.PG[tid];
  __syncthreads();
  const int n = sm.cd.n, m = a.m;
  double* Y = a.Y + sm.cd.row_off;
  const float* W = a.W + sm.cd.row_off;
  double* AF = a.AF + sm.cd.row_off;
  double* TM = a.TM + sm.cd.row_off;
  const uint8_t* YB = a.ybin + sm.cd.row_off;
  const uint8_t* Q = a.q + sm.cd.q_off;
  uint8_t* LEAF = a.leaf + sm.cd.leaf_off;
  TreeDev* TR = a.trees + sm.cd.tree_off;
  ChainState* cs = a.state + chain;
  const LibmTables tb = a.tb;
  const double t2 = (double)sm.cd.tau * sm.cd.tau;
  const double a_prior = 1 / t2;

  Rng g{0, 0, 0, 0};
  double sigma = 0;
  int n_birth = 0, n_death = 0, n_cap = 0;
  if (tid == 0) {
    g = Rng{cs->rng[0], cs->rng[1], cs->rng[2], cs->rng[3]};
    sigma = cs->sigma;
    n_birth = cs->n_birth;
    n_death = cs->n_death;
    n_cap = cs->n_cap;
  }

  for (int it = a.it0; it < a.it1; it++) {
    if (tid == 0) {
      sm.inv_s2 = 1.0 / (sigma * sigma);
      sm.pmode = 0;
    }
    __syncthreads();
    for (int j = 0; j < m; j++) {
      // ---- tree j into LDS
      {
        const uint64_t* src = reinterpret_cast<const uint64_t*>(&TR[j]);
        uint64_t* dst = reinterpret_cast<uint64_t*>(&sm.T);
        for (int w = tid; w < (int)(sizeof(TreeDev) / 8); w += kThreads) dst[w] = src[w];
      }
      __syncthreads();
      int ngood = 0, nnogs = 0;
      if (tid == 0) propose(sm, g, ngood, nnogs);
      if (tid < kNodes) {
        const int s = tid;
        sm.thc[s] = (double)sm.T.th[s];
        const uint64_t lm = sm.T.used & sm.T.leaf;
        if (BITM(lm, s)) {
          const int rk = __popcll(lm & ((1ull << s) - 1ull));
          sm.rank_of[s] = (uint8_t)rk;
          sm.slot_of[rk] = (uint8_t)s;
        } else {
          sm.rank_of[s] = 0;
        }
        if (s == 0) sm.nleaf = __popcll(lm);
      }
      __syncthreads();
      // ---- data pass: apply tree j-1's change, statistics of tree j
      data_pass(a, sm, AF, TM, Y, W, Q, LEAF, j - 1, j, j > 0, true, false, false);

      // ---- Metropolis-Hastings and the generator outputs of this update's 64 normals (:469-518)
      if (tid == 0) {
        TreeDev& t = sm.T;
        const double inv_s2 = sm.inv_s2;
        const double lu = log_unif(g, tb);
        double lalpha = 0;
        int accepted = 0, capped = 0;
        int nmode = 0, nL_ = 0, nR_ = 0, nl_ = 0, nr_ = 0;
        if (sm.cmode == 1) {
          if (sm.nL >= 5 && sm.nR >= 5) {
            const double bl = sm.sL0 * inv_s2, Ml = sm.sL1 * inv_s2, br = sm.sR0 * inv_s2, Mr = sm.sR1 * inv_s2;
            lalpha = log(sm.prr) + lh(bl, Ml, t2) + lh(br, Mr, t2) - lh(bl + br, Ml + Mr, t2);
            if (lalpha > 0) lalpha = 0;
            if (lu < lalpha) {
              const int L = first_free(t.used, -1, a.node_cap);
              const int R = L < 0 ? -1 : first_free(t.used, L, a.node_cap);
              if (L < 0 || R < 0) {
                n_cap++;
                capped = 1;
              } else {
                const int nx = sm.cnx;
                t.v[nx] = (uint8_t)sm.cv;
                t.c[nx] = (uint8_t)sm.cc;
                t.l[nx] = (uint8_t)L;
                t.r[nx] = (uint8_t)R;
                t.par[L] = t.par[R] = (uint8_t)nx;
                t.l[L] = t.r[L] = t.l[R] = t.r[R] = 0;
                t.d[L] = t.d[R] = (uint8_t)(t.d[nx] + 1);
                t.used |= (1ull << L) | (1ull << R);
                t.leaf = (t.leaf & ~(1ull << nx)) | (1ull << L) | (1ull << R);
                t.ngv[L] = (uint8_t)count_goodvars(t, L, sm.cd.nc, sm.cd.p);
                t.ngv[R] = (uint8_t)count_goodvars(t, R, sm.cd.nc, sm.cd.p);
                sm.S0[L] = sm.sL0;
                sm.S1[L] = sm.sL1;
                sm.S0[R] = sm.sR0;
                sm.S1[R] = sm.sR1;
                nmode = 1;
                nL_ = L;
                nR_ = R;
                n_birth++;
                accepted = 1;
              }
            }
          }
        } else if (sm.cmode == 2) {
          const int nx = sm.cnx, l = t.l[nx], r = t.r[nx];
          const double bl = sm.S0[l] * inv_s2, Ml = sm.S1[l] * inv_s2, br = sm.S0[r] * inv_s2,
                       Mr = sm.S1[r] * inv_s2;
          lalpha = log(sm.prr) + lh(bl + br, Ml + Mr, t2) - lh(bl, Ml, t2) - lh(br, Mr, t2);
          if (lalpha > 0) lalpha = 0;
          if (lu < lalpha) {
            sm.S0[nx] = sm.S0[l] + sm.S0[r];
            sm.S1[nx] = sm.S1[l] + sm.S1[r];
            t.used &= ~((1ull << l) | (1ull << r));
            t.leaf = (t.leaf & ~((1ull << l) | (1ull << r))) | (1ull << nx);
            t.l[nx] = t.r[nx] = t.v[nx] = t.c[nx] = 0;
            nmode = 2;
            nl_ = l;
            nr_ = r;
            n_death++;
            accepted = 1;
          }
        }
        // node_cap is the reference's NODES: that many normals per update (:521-523)
        for (int s = 0; s < 2 * a.node_cap; s++) sm.rnd[s] = rng_next(g);
        // the previous-update record for the next pass (:470, :500, :515)
        sm.pmode = nmode;
        sm.pnx = sm.cnx;
        sm.pv = sm.cv;
        sm.pc = sm.cc;
        sm.pL = nL_;
        sm.pR = nR_;
        sm.pl = nl_;
        sm.pr = nr_;
        if (a.trace) {
          const long idx = (long)(it - a.it0) * m + j;
          if (idx < a.trace_cap) {
            TraceRec& tr = a.trace[(long)chain * a.trace_cap + idx];
            tr.it = it;
            tr.j = j;
            tr.mode = sm.cmode;
            tr.nx = sm.cnx;
            tr.v = sm.cv;
            tr.c = sm.cc;
            tr.nl = sm.nL;
            tr.nr = sm.nR;
            tr.accepted = accepted;
            tr.ngood = ngood;
            tr.nnogs = nnogs;
            tr.cap = capped;
            tr.pr = sm.prr;
            tr.lalpha = lalpha;
            tr.lu = lu;
            tr.sl0 = sm.sL0;
            tr.sl1 = sm.sL1;
            tr.sr0 = sm.sR0;
            tr.sr1 = sm.sR1;
            // generator state after this update's 64 normals
            Rng h = g;
            tr.rng_after[0] = h.s0;
            tr.rng_after[1] = h.s1;
            tr.rng_after[2] = h.s2;
            tr.rng_after[3] = h.s3;
          }
        }
      }
      __syncthreads();
      // ---- leaf values: one normal per slot, used leaves keep theirs (:520-528)
      if (tid < kNodes) {
        const int s = tid;
        if (s < a.node_cap && BITM(sm.T.leaf, s) && BITM(sm.T.used, s)) {
          const float z = normal_f(sm.rnd[2 * s], sm.rnd[2 * s + 1], tb);
          const double inv_s2 = sm.inv_s2;
          const double b = sm.S0[s] * inv_s2, M = sm.S1[s] * inv_s2;
          sm.T.th[s] = (float)(M / (a_prior + b) + (double)z / __dsqrt_rn(a_prior + b));
        }
        sm.tho[s] = sm.thc[s];
      }
      __syncthreads();
      if (tid < kNodes) sm.thn[tid] = (double)sm.T.th[tid];
      {
        const uint64_t* src = reinterpret_cast<const uint64_t*>(&sm.T);
        uint64_t* dst = reinterpret_cast<uint64_t*>(&TR[j]);
        for (int w = tid; w < (int)(sizeof(TreeDev) / 8); w += kThreads) dst[w] = src[w];
      }
      __syncthreads();
    }
    // ---- closing pass (:536-541)
    const bool keep = it >= a.nskip;
    data_pass(a, sm, AF, TM, Y, W, Q, LEAF, m - 1, m, true, false, true, keep);
    if (tid == 0) {
      if (a.probit) {
        // latent redraw, rows in order (:543-547)
        const double off = (double)sm.cd.offset;
        for (int i = 0; i < n; i++) {
          const double f = AF[i];
          Y[i] = YB[i] == 0 ? -rtnorm_dev(g, -f, off, tb) : rtnorm_dev(g, f, -off, tb);
        }
      } else {
        const float nul = sm.cd.nu * sm.cd.lambda;
        const float nnu = (float)n + sm.cd.nu;
        sigma = __dsqrt_rn(((double)nul + sm.rss) / (2.0 * rgamma_dev(g, (double)nnu / 2.0, tb)));
      }
      a.sigma_out[sm.cd.sig_off + it] = (float)sigma;
      sm.pmode = 0;
    }
    if (keep && a.keep_trees) {
      uint64_t* out = a.trees_out + sm.cd.tout_off + (size_t)(it - a.nskip) * m * kNodes;
      for (int idx = tid; idx < m * kNodes; idx += kThreads) {
        const int jj = idx / kNodes, s = idx % kNodes;
        out[idx] = pack_word(TR[jj], s);
      }
    }
    __syncthreads();
  }
  if (tid == 0) {
    cs->rng[0] = g.s0;
    cs->rng[1] = g.s1;
    cs->rng[2] = g.s2;
    cs->rng[3] = g.s3;
    cs->sigma = sigma;
    cs->it = a.it1;
    cs->n_birth = n_birth;
    cs->n_death = n_death;
    cs->n_cap = n_cap;
  }
}

// Chain initialisation: root-only trees, generator seed + 16 discarded outputs, y or the initial probit latents
// (src/bart_kernel.cpp:367-386).
__global__ __launch_bounds__(kThreads) void k_init(KArgs a, const float* yin, const unsigned* seeds) {
  const int tid = threadIdx.x;
  const int chain = blockIdx.x;
  const ChainDesc cd = a.desc[chain];
  TreeDev* TR = a.trees + cd.tree_off;
  for (int w = tid; w < a.m * (int)(sizeof(TreeDev) / 8); w += kThreads)
    reinterpret_cast<uint64_t*>(TR)[w] = 0;
  double* Y = a.Y + cd.row_off;
  for (int i = tid; i < cd.n; i += kThreads) {
    Y[i] = (double)yin[cd.row_off + i];
    a.AF[cd.row_off + i] = 0;
    a.TM[cd.row_off + i] = 0;
  }
  __syncthreads();
  for (int j = tid; j < a.m; j += kThreads) {
    TR[j].ngv[0] = (uint8_t)cd.root_ngv;
    TR[j].used = 1;
    TR[j].leaf = 1;
  }
  if (tid == 0) {
    const uint32_t seed = seeds[chain];
    Rng g{seed ^ 0x9E3779B9u, seed * 2654435761u + 1u, 0x6A09E667u ^ (seed << 7), 0xBB67AE85u + seed};
    for (int i = 0; i < 16; i++) rng_next(g);
    if (a.probit) {
      const double off = (double)cd.offset;
      const uint8_t* YB = a.ybin + cd.row_off;
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
  }
}

// trmean = (float)TM / (float)ndpost (:573-578)
__global__ void k_trmean(const double* TM, float* out, int n, int ndpost) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)TM[i] / (float)ndpost;
}

}  // namespace fbh

// ==================================================================== host side
using namespace fbh;

struct fbh_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int C = 0, m = 0, nskip = 0, ndpost = 0, probit = 0, keep_trees = 0, stats_mode = 0, node_cap = kNodes;
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
  LibmTables tb{};
  size_t rows_total = 0, q_total = 0, leaf_total = 0;
  std::vector<hipEvent_t> ev;  // start/stop pairs per launch
  double bytes_per_sweep = 0;
};

static void free_batch(fbh_batch* b) {
  if (!b) return;
  if (b->device >= 0) (void)hipSetDevice(b->device);
  void* ptrs[] = {b->d_desc, b->d_state, b->d_trees, b->d_q, b->d_Y, b->d_W, b->d_yin, b->d_AF, b->d_TM,
                  b->d_ybin, b->d_leaf, b->d_sigma, b->d_tout, b->d_PG, b->d_trm, b->d_trace};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto e : b->ev) (void)hipEventDestroy(e);
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

static KArgs make_args(const fbh_batch* b) {
  KArgs a;
  a.desc = b->d_desc;
  a.state = b->d_state;
  a.trees = b->d_trees;
  a.q = b->d_q;
  a.Y = b->d_Y;
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
  return a;
}

extern "C" int fbh_batch_create(const fbh_chain_in* chains, int nchains, const fbh_opts* o, fbh_batch** out) {
  if (!out || !chains || !o || nchains < 1) {
    set_error(FBH_EINVAL, "fbh_batch_create: null argument or no chains");
    return FBH_EINVAL;
  }
  *out = nullptr;
  if (o->m < 1 || o->m > FBH_MMAX || o->nskip < 0 || o->ndpost < 0 || !o->pgtab ||
      (o->stats_mode != FBH_STATS_REF16 && o->stats_mode != FBH_STATS_FX64) ||
      (o->node_cap != 0 && (o->node_cap < 3 || o->node_cap > kNodes))) {
    set_error(FBH_EINVAL, "fbh_batch_create: bad options (m, nskip, ndpost, pgtab, stats_mode or node_cap)");
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
    for (int i = 0; i < ci.n; i++) {
      if (!std::isfinite(ci.wt[i]) || (!o->probit && !std::isfinite(ci.y[i]))) {
        set_error(FBH_EINVAL, "fbh_batch_create: non-finite y or weight in chain " + std::to_string(c));
        return FBH_EINVAL;
      }
    }
  }
  int rc = ensure_device(o->device);
  if (rc) return rc;
  fbh_batch* b = new fbh_batch();
  b->device = o->device;
  b->C = nchains;
  b->m = o->m;
  b->nskip = o->nskip;
  b->ndpost = o->ndpost;
  b->probit = o->probit ? 1 : 0;
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
  // ---- pools
  b->hdesc.resize(nchains);
  size_t qo = 0, ro = 0, lo = 0;
  const int total = b->nskip + b->ndpost;
  for (int c = 0; c < nchains; c++) {
    const fbh_chain_in& ci = chains[c];
    ChainDesc& d = b->hdesc[c];
    std::memset(&d, 0, sizeof d);
    d.n = ci.n;
    d.p = ci.p;
    d.tau = ci.tau;
    d.lambda = ci.lambda;
    d.nu = ci.nu;
    d.sigest = o->probit ? ci.sigest : ci.sigest;
    d.offset = ci.offset;
    int rn = 0;
    for (int v = 0; v < FBH_PMAX; v++) {
      d.nc[v] = v < ci.p ? ci.numcut[v] : 0;
      if (v < ci.p && d.nc[v] >= 1) rn++;
    }
    d.root_ngv = rn;
    d.q_off = qo;
    d.row_off = ro;
    d.leaf_off = lo;
    d.tree_off = (uint64_t)c * b->m;
    d.sig_off = (uint64_t)c * total;
    d.tout_off = (uint64_t)c * b->ndpost * b->m * kNodes;
    qo += (size_t)ci.p * ci.n;
    ro += ((size_t)ci.n + 63) / 64 * 64;  // keep each chain's rows 512-byte aligned
    lo += ((size_t)b->m * ci.n + 255) / 256 * 256;
    b->bytes_per_sweep += 37.0 * ci.n * (b->m + 1);
  }
  b->q_total = qo;
  b->rows_total = ro;
  b->leaf_total = lo;
  const size_t tout = b->keep_trees ? (size_t)nchains * b->ndpost * b->m * kNodes : 1;
  if ((rc = dalloc(&b->d_desc, nchains)) || (rc = dalloc(&b->d_state, nchains)) ||
      (rc = dalloc(&b->d_trees, (size_t)nchains * b->m)) || (rc = dalloc(&b->d_q, qo)) ||
      (rc = dalloc(&b->d_Y, ro)) || (rc = dalloc(&b->d_W, ro)) || (rc = dalloc(&b->d_yin, ro)) ||
      (rc = dalloc(&b->d_AF, ro)) || (rc = dalloc(&b->d_TM, ro)) || (rc = dalloc(&b->d_ybin, ro)) ||
      (rc = dalloc(&b->d_leaf, lo)) || (rc = dalloc(&b->d_sigma, (size_t)nchains * std::max(total, 1))) ||
      (rc = dalloc(&b->d_tout, tout)) || (rc = dalloc(&b->d_PG, kNodes)) || (rc = dalloc(&b->d_trm, ro))) {
    free_batch(b);
    return rc;
  }
  // ---- staging on the host, one copy per pool
  {
    std::vector<uint8_t> hq(qo), hyb(ro, 0);
    std::vector<float> hw(ro, 0.f), hy(ro, 0.f);
    for (int c = 0; c < nchains; c++) {
      const fbh_chain_in& ci = chains[c];
      const ChainDesc& d = b->hdesc[c];
      std::memcpy(hq.data() + d.q_off, ci.q, (size_t)ci.p * ci.n);
      std::memcpy(hw.data() + d.row_off, ci.wt, sizeof(float) * ci.n);
      if (!o->probit) std::memcpy(hy.data() + d.row_off, ci.y, sizeof(float) * ci.n);
      if (o->probit)
        for (int i = 0; i < ci.n; i++) hyb[d.row_off + i] = ci.ybin[i] != 0;
    }
    std::vector<double> pg(kNodes);
    for (int d = 0; d < kNodes; d++) pg[d] = (double)o->pgtab[d];
    std::vector<unsigned> seeds(nchains);
    for (int c = 0; c < nchains; c++) seeds[c] = chains[c].seed;
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
    if (e == hipSuccess) e = hipMemsetAsync(b->d_leaf, 0, lo, b->stream);
    if (e == hipSuccess) {
      KArgs a = make_args(b);
      hipLaunchKernelGGL(k_init, dim3(nchains), dim3(kThreads), 0, b->stream, a, b->d_yin, d_seeds);
      e = hipGetLastError();
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
  int end = std::min(total, b->it + nsweeps);
  KArgs a = make_args(b);
  while (b->it < end) {
    const int it1 = std::min(end, b->it + b->sweeps_per_launch);
    a.it0 = b->it;
    a.it1 = it1;
    hipEvent_t e0, e1;
    FBH_HIP_CHECK(hipEventCreate(&e0));
    FBH_HIP_CHECK(hipEventCreate(&e1));
    b->ev.push_back(e0);
    b->ev.push_back(e1);
    FBH_HIP_CHECK(hipEventRecord(e0, b->stream));
    hipLaunchKernelGGL(k_sweeps, dim3(b->C), dim3(kThreads), 0, b->stream, a);
    FBH_HIP_CHECK(hipGetLastError());
    FBH_HIP_CHECK(hipEventRecord(e1, b->stream));
    b->it = it1;
  }
  return FBH_OK;
}

extern "C" int fbh_batch_sync(fbh_batch* b) {
  if (!b) return FBH_EINVAL;
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  return FBH_OK;
}

extern "C" int fbh_batch_iter(const fbh_batch* b) { return b ? b->it : -1; }

extern "C" int fbh_batch_fetch(fbh_batch* b, int c, const fbh_chain_out* out) {
  if (!b || !out || c < 0 || c >= b->C) {
    set_error(FBH_EINVAL, "fbh_batch_fetch: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  const ChainDesc& d = b->hdesc[c];
  const int total = b->nskip + b->ndpost;
  if (out->sigma) FBH_HIP_CHECK(hipMemcpy(out->sigma, b->d_sigma + d.sig_off, sizeof(float) * total, hipMemcpyDeviceToHost));
  if (out->trmean) {
    hipLaunchKernelGGL(k_trmean, dim3((d.n + 255) / 256), dim3(256), 0, b->stream, b->d_TM + d.row_off,
                       b->d_trm + d.row_off, d.n, b->ndpost);
    FBH_HIP_CHECK(hipGetLastError());
    FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
    FBH_HIP_CHECK(hipMemcpy(out->trmean, b->d_trm + d.row_off, sizeof(float) * d.n, hipMemcpyDeviceToHost));
  }
  if (out->trees && b->keep_trees)
    FBH_HIP_CHECK(hipMemcpy(out->trees, b->d_tout + d.tout_off, sizeof(uint64_t) * b->ndpost * b->m * kNodes,
                            hipMemcpyDeviceToHost));
  if (out->stats) {
    ChainState cs;
    FBH_HIP_CHECK(hipMemcpy(&cs, b->d_state + c, sizeof cs, hipMemcpyDeviceToHost));
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
  if (!b || c < 0 || c >= b->C) {
    set_error(FBH_EINVAL, "fbh_batch_state: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  const ChainDesc& d = b->hdesc[c];
  ChainState cs;
  FBH_HIP_CHECK(hipMemcpy(&cs, b->d_state + c, sizeof cs, hipMemcpyDeviceToHost));
  if (rng) std::memcpy(rng, cs.rng, 16);
  if (sigma) *sigma = cs.sigma;
  if (af) FBH_HIP_CHECK(hipMemcpy(af, b->d_AF + d.row_off, sizeof(double) * d.n, hipMemcpyDeviceToHost));
  if (latent) FBH_HIP_CHECK(hipMemcpy(latent, b->d_Y + d.row_off, sizeof(double) * d.n, hipMemcpyDeviceToHost));
  if (leaf) FBH_HIP_CHECK(hipMemcpy(leaf, b->d_leaf + d.leaf_off, (size_t)b->m * d.n, hipMemcpyDeviceToHost));
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
  for (size_t i = 0; i + 1 < b->ev.size(); i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]) == hipSuccess) {
      tot += ms;
      k++;
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
    FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
    FBH_HIP_CHECK(hipMemcpy(host_out, b->d_trace + (size_t)chain * b->trace_cap, sizeof(TraceRec) * b->trace_cap,
                            hipMemcpyDeviceToHost));
  }
  return FBH_OK;
}

extern "C" void fbh_batch_destroy(fbh_batch* b) { free_batch(b); }

// the device yhat walk (fbh_predict.hip)
namespace fbh {
int yhat_device(const uint64_t* d_trees, const uint8_t* d_q, int n, int ndpost, int m, hipStream_t stream,
                double* host_out);
}

extern "C" int fbh_batch_yhat(fbh_batch* b, int c, double* yhat) {
  if (!b || c < 0 || c >= b->C || !yhat || !b->keep_trees) {
    set_error(FBH_EINVAL, "fbh_batch_yhat: bad argument or trees not kept");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(b->device));
  FBH_HIP_CHECK(hipStreamSynchronize(b->stream));
  const ChainDesc& d = b->hdesc[c];
  return fbh::yhat_device(b->d_tout + d.tout_off, b->d_q + d.q_off, d.n, b->ndpost, b->m, b->stream, yhat);
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
