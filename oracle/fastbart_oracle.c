/* fastbart_oracle.c -- TEST INFRASTRUCTURE ONLY: CPU restatement of the reference sampler, glue and predictor.
 * Only tests/, the smoke test and the benchmark's CPU leg load this file's library; the product never does.
 * Each function names the reference lines it restates (paths relative to the reference root). One part has no
 * reference counterpart: the informed birth rule of the proposal study (off unless orc_chain_set_informed is called),
 * pinned by the enumerated single-tree posterior (tests/test_tree_posterior.py).
 * Build: oracle/Makefile (-O2 -ffp-contract=off, no -march: no fused multiply-add anywhere).
 */
#include "fastbart_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ generator (src/bart_kernel.cpp:43-83) */
static inline uint32_t rotl32(uint32_t x, int k) { return (x << k) | (x >> (32 - k)); }

/* xoshiro128** step, Blackman & Vigna; same state order as src/bart_kernel.cpp:54-58 */
uint32_t orc_rng_next(orc_rng* g) {
  uint32_t* s = g->s;
  const uint32_t out = rotl32(s[1] * 5u, 7) * 9u;
  const uint32_t t = s[1] << 9;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl32(s[3], 11);
  return out;
}

/* seeding constants and the 16 discarded outputs, src/bart_kernel.cpp:381-382 */
void orc_rng_seed(orc_rng* g, uint32_t seed) {
  g->s[0] = seed ^ 0x9E3779B9u;
  g->s[1] = seed * 2654435761u + 1u;
  g->s[2] = 0x6A09E667u ^ (seed << 7);
  g->s[3] = 0xBB67AE85u + seed;
  for (int i = 0; i < 16; i++) orc_rng_next(g);
}

/* top 24 bits, centred in their cell: (k + 1/2) 2^-24, exact in float, src/bart_kernel.cpp:59 */
float orc_unif(orc_rng* g) { return ((float)(orc_rng_next(g) >> 8) + 0.5f) * (1.0f / 16777216.0f); }

/* float Box-Muller, cosine branch only, two uniforms in order, src/bart_kernel.cpp:60 */
float orc_rnorm(orc_rng* g) {
  const float u1 = orc_unif(g);
  const float u2 = orc_unif(g);
  return sqrtf(-2.f * logf(u1)) * cosf(6.28318531f * u2);
}

/* ------------------------------------------------------------------ log of the libm comparisons
 * For measuring how often a last-bit difference between two libm implementations could change a decision: every
 * accept test of rtnorm's exponential branch (u > exp(arg), src/bart_kernel.cpp:289) and of the gamma draw
 * (log u < A + d log v, :80) is seen; the first `sample` of them, and every one whose two sides are within 1e-13
 * (relative) of each other, are recorded as 7 doubles: kind (0 rtnorm, 1 gamma), the libm argument (arg or v), the
 * comparison's left side (u, or log u), A, d (gamma), the host libm's value (exp(arg) or log v), near (1 if the sides
 * are that close). Not thread-safe; test infrastructure. */
static double* g_cmp = NULL;
static long g_cmp_cap = 0, g_cmp_n = 0, g_cmp_seen = 0, g_cmp_sample = 0;

void orc_cmp_log(double* buf, long cap, long sample) {
  g_cmp = buf;
  g_cmp_cap = cap;
  g_cmp_sample = sample;
  g_cmp_n = 0;
  g_cmp_seen = 0;
}

long orc_cmp_count(long* seen) {
  if (seen) *seen = g_cmp_seen;
  return g_cmp_n;
}

static void cmp_record(int kind, double arg, double lhs, double A, double d, double hostval, double rhs) {
  if (!g_cmp) return;
  g_cmp_seen++;
  const int near = fabs(lhs - rhs) <= 1e-13 * (fabs(rhs) > 1.0 ? fabs(rhs) : 1.0);
  if ((g_cmp_seen <= g_cmp_sample || near) && g_cmp_n < g_cmp_cap) {
    double* r = g_cmp + 7 * g_cmp_n++;
    r[0] = kind;
    r[1] = arg;
    r[2] = lhs;
    r[3] = A;
    r[4] = d;
    r[5] = hostval;
    r[6] = near;
  }
}

/* Marsaglia-Tsang (2000) for shape a >= 1, at most 32 tries, src/bart_kernel.cpp:73-83 */
double orc_rgamma(orc_rng* g, double a) {
  const double d = a - 1.0 / 3.0;
  const double c = 1.0 / sqrt(9.0 * d);
  for (int tries = 0; tries < 32; tries++) {
    const double x = (double)orc_rnorm(g);
    double v = 1.0 + c * x;
    if (v <= 0) continue;
    v = v * v * v;
    const double u = (double)orc_unif(g);
    const double rhs = 0.5 * x * x + d - d * v + d * log(v);
    if (g_cmp) cmp_record(1, v, log(u), 0.5 * x * x + d - d * v, d, log(v), rhs);
    if (log(u) < rhs) return d * v;
  }
  return d;
}

/* N(mean,1) truncated to values above tau (Robert 1995 exponential proposal when the bound is above the mean),
 * src/bart_kernel.cpp:282-292 with the internal generator */
double orc_rtnorm(orc_rng* g, double mean, double tau) {
  double z;
  tau = tau - mean;
  if (tau <= 0.) {
    do {
      z = (double)orc_rnorm(g);
    } while (z < tau);
  } else {
    const double lam = 0.5 * (tau + sqrt(tau * tau + 4.0));
    for (;;) {
      const double e = -log((double)orc_unif(g));
      z = e / lam + tau;
      const double u = (double)orc_unif(g);
      if (g_cmp) {
        const double arg = -0.5 * (z - lam) * (z - lam);
        cmp_record(0, arg, u, 0.0, 0.0, exp(arg), exp(arg));
      }
      if (!(u > exp(-0.5 * (z - lam) * (z - lam)))) break;
    }
  }
  return z + mean;
}

/* ------------------------------------------------------------------ trees (src/bart_kernel.cpp:85-116, 221-276) */
typedef struct otree {
  uint8_t l[ORC_NODES], r[ORC_NODES], par[ORC_NODES], v[ORC_NODES], c[ORC_NODES], d[ORC_NODES], ngv[ORC_NODES];
  float th[ORC_NODES];
  uint64_t used, leaf;
} otree;

#define BIT(mask, i) (((mask) >> (i)) & 1ull)

static int popc64(uint64_t m) { return __builtin_popcountll(m); }

/* index of the k-th set bit (0-based); 0 when there are not k+1 bits, src/bart_kernel.cpp:94-101 */
static int select_kth(uint64_t m, int k) {
  int seen = 0;
  for (int i = 0; i < ORC_NODES; i++)
    if (BIT(m, i)) {
      if (seen == k) return i;
      seen++;
    }
  return 0;
}

/* lowest unused slot in [1, cap) other than skip, or -1, src/bart_kernel.cpp:102-109 */
static int first_free(uint64_t used, int skip, int cap) {
  for (int i = 1; i < cap; i++)
    if (!BIT(used, i) && i != skip) return i;
  return -1;
}

/* cutpoint range [L,U] of variable v at node x from the ancestors' rules, src/bart_kernel.cpp:238-251 */
static void node_range(const otree* t, int x, int v, int nc, int* L, int* U) {
  int lo = 0, hi = nc - 1, node = x;
  for (int s = 0; s < ORC_MAXDEPTH && node != 0; s++) {
    const int pa = t->par[node];
    if ((int)t->v[pa] == v) {
      if ((int)t->l[pa] == node) {
        if ((int)t->c[pa] - 1 < hi) hi = (int)t->c[pa] - 1;
      } else {
        if ((int)t->c[pa] + 1 > lo) lo = (int)t->c[pa] + 1;
      }
    }
    node = pa;
  }
  *L = lo;
  *U = hi;
}

/* number of variables with a non-empty range at node x, src/bart_kernel.cpp:253-276 */
static int count_goodvars(const otree* t, int x, const int* nc, int p) {
  int lo[ORC_PMAX], hi[ORC_PMAX];
  for (int v = 0; v < ORC_PMAX; v++) {
    lo[v] = 0;
    hi[v] = nc[v] - 1;
  }
  int node = x;
  for (int s = 0; s < ORC_MAXDEPTH && node != 0; s++) {
    const int pa = t->par[node], v = t->v[pa], c = t->c[pa];
    if ((int)t->l[pa] == node) {
      if (c - 1 < hi[v]) hi[v] = c - 1;
    } else {
      if (c + 1 > lo[v]) lo[v] = c + 1;
    }
    node = pa;
  }
  int g = 0;
  for (int v = 0; v < p; v++) g += hi[v] >= lo[v];
  return g;
}

/* log integrated likelihood of one leaf, src/bart_kernel.cpp:278 */
static double lh(double b, double M, double t2) {
  const double k = b * t2 + 1;
  return -0.5 * log(k) + 0.5 * M * M * t2 / k;
}

/* output word, src/bart_kernel.cpp:565-569 */
static uint64_t pack_slot(const otree* t, int s) {
  uint32_t thb;
  memcpy(&thb, &t->th[s], 4);
  const uint64_t vv = BIT(t->used, s) ? (uint64_t)t->v[s] : 255u;
  return vv | ((uint64_t)t->c[s] << 8) | ((uint64_t)t->l[s] << 16) | ((uint64_t)t->r[s] << 24) | ((uint64_t)thb << 32);
}

/* ------------------------------------------------------------------ chain */
typedef struct oprev {
  int mode, nx, v, c, L, R, l, r;
  double tho[ORC_NODES], thn[ORC_NODES];
} oprev;
typedef struct ocur {
  int mode, nx, v, c;
  double th[ORC_NODES];
} ocur;
typedef struct ostats {
  double S0[ORC_NODES], S1[ORC_NODES], sL0, sL1, sR0, sR1, rss;
  int nL, nR;
} ostats;

struct orc_chain {
  int n, p, m, nskip, ndpost, probit, node_cap, stats_mode;
  float tau, lambda, nu, sigest, offset;
  int nc[ORC_PMAX];
  double PG[ORC_NODES];
  uint8_t* q;    /* [p][n] */
  double* Y;     /* [n] y, or the latent z for probit */
  double* W;     /* [n] */
  double* AF;    /* [n] running fit */
  double* TM;    /* [n] sum of the fit over kept sweeps */
  int* ybin;     /* [n] probit outcomes */
  uint8_t* leaf; /* [m][n] */
  otree* T;      /* [m] */
  orc_rng rng;
  double sigma, t2, a_prior;
  int it;
  int n_birth, n_death, n_cap;
  orc_trace* trace;
  long trace_cap, trace_len;
  double informed_alpha; /* proposal study (no reference counterpart): 0 = the reference's proposal */
  long informed_births;  /* births proposed from the informed rule */
};

static inline int64_t fx_round(double x, int shift) { return (int64_t)llrint(ldexp(x, shift)); }

/* One pass over the rows: apply tree jp's accepted change (leaf ids and running fit), then the sufficient statistics of
 * tree j's proposal. src/bart_kernel.cpp:131-213. REF16 keeps 16 double lanes (row i in lane i%16, each lane summed in
 * row order, lanes added 0..15 from zero); FX64 sums int64 fixed-point images of the same per-row values. */
static void data_pass(orc_chain* ch, int jp, int j, const oprev* pu, int apply, const ocur* cp, int stats, int rss,
                      int keep, ostats* st) {
  const int n = ch->n;
  static __thread double a0[ORC_KL][ORC_NODES], a1[ORC_KL][ORC_NODES];
  double bl0[ORC_KL], bl1[ORC_KL], br0[ORC_KL], br1[ORC_KL], ra[ORC_KL];
  int nl[ORC_KL], nr[ORC_KL];
  static __thread int64_t f0[ORC_NODES], f1[ORC_NODES];
  int64_t fl0 = 0, fl1 = 0, fr0 = 0, fr1 = 0, frss = 0;
  memset(a0, 0, sizeof a0);
  memset(a1, 0, sizeof a1);
  memset(f0, 0, sizeof f0);
  memset(f1, 0, sizeof f1);
  for (int k = 0; k < ORC_KL; k++) {
    bl0[k] = bl1[k] = br0[k] = br1[k] = ra[k] = 0;
    nl[k] = nr[k] = 0;
  }
  const int fx = ch->stats_mode == ORC_STATS_FX64;
  uint8_t* lprev = apply ? ch->leaf + (size_t)jp * n : NULL;
  const uint8_t* lcur = stats ? ch->leaf + (size_t)j * n : NULL;
  const uint8_t* qp = apply && pu->mode == 1 ? ch->q + (size_t)pu->v * n : NULL;
  const uint8_t* qc = stats && cp->mode == 1 ? ch->q + (size_t)cp->v * n : NULL;
  for (int i = 0; i < n; i++) {
    const int k = i % ORC_KL;
    double af = ch->AF[i];
    if (apply) {
      const int lf = lprev[i];
      int lf2 = lf;
      if (pu->mode == 1 && lf == pu->nx)
        lf2 = qp[i] <= pu->c ? pu->L : pu->R;
      else if (pu->mode == 2 && (lf == pu->l || lf == pu->r))
        lf2 = pu->nx;
      af = af - pu->tho[lf] + pu->thn[lf2];
      lprev[i] = (uint8_t)lf2;
      ch->AF[i] = af;
    }
    if (rss) {
      const double e = ch->Y[i] - af;
      const double c2 = (ch->W[i] * e) * e;
      if (fx)
        frss += fx_round(c2, ORC_FX_RSS_SHIFT);
      else
        ra[k] += c2;
      if (keep) ch->TM[i] += af;
    }
    if (stats) {
      const int lf = lcur[i];
      const double r = ch->Y[i] - af + cp->th[lf];
      const double w = ch->W[i];
      const double wr = w * r;
      if (fx) {
        const int64_t iw = fx_round(w, ORC_FX_SHIFT), iwr = fx_round(wr, ORC_FX_SHIFT);
        f0[lf] += iw;
        f1[lf] += iwr;
        if (cp->mode == 1 && lf == cp->nx) {
          if (qc[i] <= cp->c) {
            nl[k]++;
            fl0 += iw;
            fl1 += iwr;
          } else {
            nr[k]++;
            fr0 += iw;
            fr1 += iwr;
          }
        }
      } else {
        a0[k][lf] += w;
        a1[k][lf] += wr;
        if (cp->mode == 1 && lf == cp->nx) {
          if (qc[i] <= cp->c) {
            nl[k]++;
            bl0[k] += w;
            bl1[k] += wr;
          } else {
            nr[k]++;
            br0[k] += w;
            br1[k] += wr;
          }
        }
      }
    }
  }
  st->nL = st->nR = 0;
  for (int k = 0; k < ORC_KL; k++) {
    st->nL += nl[k];
    st->nR += nr[k];
  }
  if (fx) {
    for (int s = 0; s < ORC_NODES; s++) {
      st->S0[s] = ldexp((double)f0[s], -ORC_FX_SHIFT);
      st->S1[s] = ldexp((double)f1[s], -ORC_FX_SHIFT);
    }
    st->sL0 = ldexp((double)fl0, -ORC_FX_SHIFT);
    st->sL1 = ldexp((double)fl1, -ORC_FX_SHIFT);
    st->sR0 = ldexp((double)fr0, -ORC_FX_SHIFT);
    st->sR1 = ldexp((double)fr1, -ORC_FX_SHIFT);
    st->rss = ldexp((double)frss, -ORC_FX_RSS_SHIFT);
    return;
  }
  for (int s = 0; s < ORC_NODES; s++) {
    double x0 = 0, x1 = 0;
    for (int k = 0; k < ORC_KL; k++) {
      x0 += a0[k][s];
      x1 += a1[k][s];
    }
    st->S0[s] = x0;
    st->S1[s] = x1;
  }
  st->sL0 = st->sL1 = st->sR0 = st->sR1 = st->rss = 0;
  for (int k = 0; k < ORC_KL; k++) {
    st->sL0 += bl0[k];
    st->sL1 += bl1[k];
    st->sR0 += br0[k];
    st->sR1 += br1[k];
    st->rss += ra[k];
  }
}

orc_chain* orc_chain_new(const uint8_t* q, const float* y, const float* wt, const int* numcut, const float* pgtab,
                         int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
                         unsigned seed, int probit, const int* ybin, float offset, const orc_opts* opts) {
  if (n < 1 || n > ORC_NMAX || p < 1 || p > ORC_PMAX || m < 1 || m > ORC_MMAX || nskip < 0 || ndpost < 0) return NULL;
  if (probit && !ybin) return NULL;
  orc_chain* ch = (orc_chain*)calloc(1, sizeof *ch);
  if (!ch) return NULL;
  ch->n = n;
  ch->p = p;
  ch->m = m;
  ch->nskip = nskip;
  ch->ndpost = ndpost;
  ch->probit = probit;
  ch->tau = tau;
  ch->lambda = lambda;
  ch->nu = nu;
  ch->sigest = sigest;
  ch->offset = offset;
  ch->stats_mode = opts ? opts->stats_mode : ORC_STATS_REF16;
  ch->node_cap = (opts && opts->node_cap) ? opts->node_cap : ORC_NODES;
  if (ch->node_cap < 3 || ch->node_cap > ORC_NODES) ch->node_cap = ORC_NODES;
  ch->trace = opts ? opts->trace : NULL;
  ch->trace_cap = opts ? opts->trace_cap : 0;
  for (int v = 0; v < ORC_PMAX; v++) ch->nc[v] = v < p ? numcut[v] : 0;
  for (int d = 0; d < ORC_NODES; d++) ch->PG[d] = (double)pgtab[d];
  ch->q = (uint8_t*)malloc((size_t)p * n);
  ch->Y = (double*)malloc(sizeof(double) * n);
  ch->W = (double*)malloc(sizeof(double) * n);
  ch->AF = (double*)calloc(n, sizeof(double));
  ch->TM = (double*)calloc(n, sizeof(double));
  ch->leaf = (uint8_t*)calloc((size_t)m * n, 1);
  ch->T = (otree*)calloc(m, sizeof(otree));
  ch->ybin = probit ? (int*)malloc(sizeof(int) * n) : NULL;
  if (!ch->q || !ch->Y || !ch->W || !ch->AF || !ch->TM || !ch->leaf || !ch->T || (probit && !ch->ybin)) {
    orc_chain_free(ch);
    return NULL;
  }
  memcpy(ch->q, q, (size_t)p * n);
  for (int i = 0; i < n; i++) {
    ch->Y[i] = (double)y[i];
    ch->W[i] = (double)wt[i];
  }
  if (probit) memcpy(ch->ybin, ybin, sizeof(int) * n);
  /* root-only trees; the root's good-variable count is the number of variables with a cutpoint (:372-379) */
  int root_ngv = 0;
  for (int v = 0; v < p; v++) root_ngv += ch->nc[v] >= 1;
  for (int j = 0; j < m; j++) {
    ch->T[j].ngv[0] = (uint8_t)root_ngv;
    ch->T[j].used = 1;
    ch->T[j].leaf = 1;
  }
  orc_rng_seed(&ch->rng, seed);
  /* initial latents, rows in order (:384-385) */
  if (probit)
    for (int i = 0; i < n; i++)
      ch->Y[i] = ybin[i] == 0 ? -orc_rtnorm(&ch->rng, 0., (double)offset) : orc_rtnorm(&ch->rng, 0., -(double)offset);
  ch->t2 = (double)tau * tau;
  ch->a_prior = 1 / ch->t2;
  ch->sigma = sigest;
  return ch;
}

void orc_chain_free(orc_chain* ch) {
  if (!ch) return;
  free(ch->q);
  free(ch->Y);
  free(ch->W);
  free(ch->AF);
  free(ch->TM);
  free(ch->leaf);
  free(ch->T);
  free(ch->ybin);
  free(ch);
}

/* ------------------------------------------------------------------ informed birth rule (no reference counterpart)
 * A proposal study, off unless orc_chain_set_informed sets alpha > 0. A birth at node nx draws its rule (v, c) with
 * probability alpha from the rule's conditional given the tree, sigma and the partial residuals of the tree's rows,
 * and otherwise as the reference does; the mixture's probability of the rule replaces the uniform one in the
 * Metropolis-Hastings ratio of births and of deaths (the reverse of a death is a birth at the merged node with its old
 * rule), so the chain's target is the reference's. The conditional is
 *   q(v, c) proportional to (1 - PGl)(1 - PGr) PD / (U_v - L_v + 1) * exp(lh(left) + lh(right))
 * over the rules with a non-empty range whose children hold at least 5 rows each (the ratio refuses smaller ones);
 * the other factors of the birth ratio do not depend on the rule. It is the best a birth proposal learned from node
 * features can approach. Sums are plain sequential doubles; the forward and the reverse evaluation see the same rows
 * and residuals, equal up to the rounding of the running fit. */
typedef struct orule {
  int v, c;
  double lw;
} orule;

/* The rules of the conditional and their log weights at one node: rows with member[i] != 0, weights w, partial
 * residuals r (q [p][n] cutpoint ranks; ranges[u] = L | U << 8, U < L: no cut). Sums are plain sequential doubles in
 * row order. Returns the number of rules written to rules[] (variables in order, cuts ascending) and their largest log
 * weight in *lmax. */
static int split_rules(int n, int p, const uint8_t* q, const uint8_t* member, const double* w, const double* r,
                       const uint16_t* ranges, int d, int ngv, int ngood_pre, const double* PG, double inv_s2,
                       double t2, orule* rules, double* lmax_out) {
  static __thread double H0[ORC_PMAX][256], H1[ORC_PMAX][256];
  static __thread int Hn[ORC_PMAX][256];
  const double pb = 0.5;
  for (int u = 0; u < p; u++) {
    memset(H0[u], 0, sizeof H0[u]);
    memset(H1[u], 0, sizeof H1[u]);
    memset(Hn[u], 0, sizeof Hn[u]);
  }
  double B0 = 0, B1 = 0;
  int N = 0;
  for (int i = 0; i < n; i++) {
    if (!member[i]) continue;
    const double wi = w[i], wr = wi * r[i];
    B0 += wi;
    B1 += wr;
    N++;
    for (int u = 0; u < p; u++) {
      const int k = q[(size_t)u * n + i];
      H0[u][k] += wi;
      H1[u][k] += wr;
      Hn[u][k]++;
    }
  }
  int nr = 0;
  double lmax = -INFINITY;
  for (int u = 0; u < p; u++) {
    const int L = ranges[u] & 255, U = ranges[u] >> 8;
    if (U < L) continue;
    double c0 = 0, c1 = 0;
    int cn = 0;
    for (int k = 0; k <= U; k++) {
      c0 += H0[u][k];
      c1 += H1[u][k];
      cn += Hn[u][k];
      if (k < L || cn < 5 || N - cn < 5) continue;
      double PGl, PGr;
      if (ngv > 1) {
        PGl = PGr = PG[d + 1];
      } else {
        PGl = (k - 1 < L) ? 0 : PG[d + 1];
        PGr = (U < k + 1) ? 0 : PG[d + 1];
      }
      const double PD = (ngood_pre > 1) ? 1 - pb : ((PGr == 0 && PGl == 0) ? 1 : 1 - pb);
      const double lw = log((1 - PGl) * (1 - PGr) * PD / (U - L + 1)) + lh(c0 * inv_s2, c1 * inv_s2, t2) +
                        lh((B0 - c0) * inv_s2, (B1 - c1) * inv_s2, t2);
      rules[nr].v = u;
      rules[nr].c = k;
      rules[nr].lw = lw;
      nr++;
      if (lw > lmax) lmax = lw;
    }
  }
  *lmax_out = lmax;
  return nr;
}

int orc_split_conditional(int n, int p, const uint8_t* q, const uint8_t* member, const double* w, const double* r,
                          const uint16_t* ranges, int depth, int ngv, int ngood, const double* pg, double sigma,
                          double tau, double* q_out, double* entropy) {
  static __thread orule rules[ORC_PMAX * 256];
  if (n < 1 || p < 1 || p > ORC_PMAX || depth < 0 || depth + 1 >= ORC_NODES) return -1;
  double lmax;
  const int nr = split_rules(n, p, q, member, w, r, ranges, depth, ngv, ngood, pg, 1.0 / (sigma * sigma), tau * tau,
                             rules, &lmax);
  memset(q_out, 0, sizeof(double) * (size_t)p * 255);
  *entropy = -1;
  if (nr == 0) return 0;
  double Z = 0, E = 0;
  for (int k = 0; k < nr; k++) {
    const double x = rules[k].lw - lmax, e = exp(x);
    Z += e;
    E += e * x;
  }
  for (int k = 0; k < nr; k++) q_out[(size_t)rules[k].v * 255 + rules[k].c] = exp(rules[k].lw - lmax) / Z;
  *entropy = log(Z) - E / Z;
  return nr;
}

/* the conditional at node nx over the rows whose leaf in tree j is nx (or l or r: the merged node of a death);
 * draw = 1: draw a rule into *v, *c and return its probability (0 if no rule qualifies); draw = 0: return the
 * probability of rule (*v, *c) */
static double informed_rule(orc_chain* ch, const otree* t, int j, int nx, int l, int r, const double* th, double inv_s2,
                            int ngood_pre, int draw, int* v, int* c) {
  static __thread orule rules[ORC_PMAX * 256];
  static __thread uint8_t* mem = NULL;
  static __thread double* res = NULL;
  static __thread int cap = 0;
  const int n = ch->n, p = ch->p;
  if ((int)t->d[nx] + 1 >= ORC_MAXDEPTH) return 0;
  if (cap < n) {
    free(mem);
    free(res);
    mem = (uint8_t*)malloc((size_t)n);
    res = (double*)malloc(sizeof(double) * (size_t)n);
    cap = n;
  }
  const uint8_t* lf = ch->leaf + (size_t)j * n;
  for (int i = 0; i < n; i++) {
    const int s = lf[i];
    mem[i] = (uint8_t)(l < 0 ? s == nx : (s == l || s == r));
    res[i] = mem[i] ? ch->Y[i] - ch->AF[i] + th[s] : 0;
  }
  uint16_t ranges[ORC_PMAX];
  for (int u = 0; u < p; u++) {
    int L, U;
    node_range(t, nx, u, ch->nc[u], &L, &U);
    ranges[u] = U < L ? (uint16_t)0x00FF : (uint16_t)(L | (U << 8));
  }
  double lmax;
  const int nr = split_rules(n, p, ch->q, mem, ch->W, res, ranges, t->d[nx], count_goodvars(t, nx, ch->nc, p),
                             ngood_pre, ch->PG, inv_s2, ch->t2, rules, &lmax);
  if (nr == 0) return 0;
  double Z = 0;
  for (int k = 0; k < nr; k++) Z += exp(rules[k].lw - lmax);
  if (draw) {
    const double u = (double)orc_unif(&ch->rng) * Z;
    double acc = 0;
    int k = 0;
    for (; k < nr - 1; k++) {
      acc += exp(rules[k].lw - lmax);
      if (u < acc) break;
    }
    *v = rules[k].v;
    *c = rules[k].c;
    return exp(rules[k].lw - lmax) / Z;
  }
  for (int k = 0; k < nr; k++)
    if (rules[k].v == *v && rules[k].c == *c) return exp(rules[k].lw - lmax) / Z;
  return 0;
}

void orc_chain_set_informed(orc_chain* ch, double alpha) {
  ch->informed_alpha = alpha > 0 ? (alpha < 1 ? alpha : 1) : 0;
}
long orc_chain_informed_births(const orc_chain* ch) { return ch->informed_births; }

/* One sweep: m tree updates then the closing pass and the sigma / latent draw. src/bart_kernel.cpp:396-572 */
static void one_sweep(orc_chain* ch, float* sigma_out, uint64_t* trees_out, int keep_trees) {
  const int p = ch->p, m = ch->m, it = ch->it;
  const double pb = 0.5, t2 = ch->t2;
  const double inv_s2 = 1.0 / (ch->sigma * ch->sigma);
  oprev pu;
  memset(&pu, 0, sizeof pu);
  ostats st;
  const double ia = ch->informed_alpha;
  for (int j = 0; j < m; j++) {
    otree* t = &ch->T[j];
    /* the informed rule reads tree j's partial residuals before the pass: apply tree j-1's change first */
    const int pre = ia > 0 && j > 0;
    if (pre) {
      ocur none;
      memset(&none, 0, sizeof none);
      data_pass(ch, j - 1, j, &pu, 1, &none, 0, 0, 0, &st);
    }
    /* ---- proposal, :413-464 */
    uint64_t goodm = 0, nogm = 0;
    for (int s = 0; s < ORC_NODES; s++) {
      const int isleaf = (int)BIT(t->leaf, s), isused = (int)BIT(t->used, s);
      if (isused && isleaf && t->ngv[s] > 0) goodm |= 1ull << s;
      if (isused && !isleaf && BIT(t->leaf, t->l[s]) && BIT(t->leaf, t->r[s])) nogm |= 1ull << s;
    }
    const int ngood = popc64(goodm), nnogs = popc64(nogm);
    const int single = t->used == 1ull;
    const double PBx = ngood == 0 ? 0 : (single ? 1 : pb);
    ocur cp;
    cp.mode = 0;
    cp.nx = cp.v = cp.c = 0;
    for (int s = 0; s < ORC_NODES; s++) cp.th[s] = (double)t->th[s];
    double pr = 0;
    if ((double)orc_unif(&ch->rng) < PBx) {
      const int nx = select_kth(goodm, (int)((double)orc_unif(&ch->rng) * ngood));
      int v = 0, c = 0, bL = 0, bU = -1, inf = 0;
      double qx = 0;
      if (ia > 0 && (double)orc_unif(&ch->rng) < ia) {
        inf = 1;
        qx = informed_rule(ch, t, j, nx, -1, -1, cp.th, inv_s2, ngood, 1, &v, &c);
        if (qx > 0) node_range(t, nx, v, ch->nc[v], &bL, &bU);
      } else {
        v = (int)((double)orc_unif(&ch->rng) * p);
        if (v >= p) v = p - 1;
        node_range(t, nx, v, ch->nc[v], &bL, &bU);
      }
      if (bU >= bL && (int)t->d[nx] + 1 < ORC_MAXDEPTH) {
        if (!inf) {
          c = bL + (int)((double)orc_unif(&ch->rng) * (bU - bL + 1));
          if (c > bU) c = bU;
        }
        const int dnx = t->d[nx];
        const double PGnx = ch->PG[dnx];
        double PGly, PGry;
        if (t->ngv[nx] > 1) {
          PGly = PGry = ch->PG[dnx + 1];
        } else {
          PGly = (c - 1 < bL) ? 0 : ch->PG[dnx + 1];
          PGry = (bU < c + 1) ? 0 : ch->PG[dnx + 1];
        }
        const double PDy = (ngood > 1) ? 1 - pb : ((PGry == 0 && PGly == 0) ? 1 : 1 - pb);
        double Pnogy;
        if (nx == 0) {
          Pnogy = 1;
        } else {
          const int pa = t->par[nx];
          const int panog = BIT(t->leaf, t->l[pa]) && BIT(t->leaf, t->r[pa]);
          Pnogy = panog ? 1.0 / nnogs : 1.0 / (nnogs + 1);
        }
        const double Pbotx = 1.0 / ngood;
        pr = (PGnx * (1 - PGly) * (1 - PGry) * PDy * Pnogy) / ((1 - PGnx) * Pbotx * PBx);
        if (ia > 0) {
          const double qu = (1.0 / p) / (bU - bL + 1);
          if (!inf) qx = informed_rule(ch, t, j, nx, -1, -1, cp.th, inv_s2, ngood, 0, &v, &c);
          pr *= qu / ((1 - ia) * qu + ia * qx);
          ch->informed_births += inf;
        }
        cp.mode = 1;
        cp.nx = nx;
        cp.v = v;
        cp.c = c;
      }
    } else {
      const int nx = select_kth(nogm, (int)((double)orc_unif(&ch->rng) * nnogs));
      const int dny = t->d[nx], l = t->l[nx], r = t->r[nx];
      const double PGny = ch->PG[dny];
      const double PGlx = t->ngv[l] > 0 ? ch->PG[dny + 1] : 0, PGrx = t->ngv[r] > 0 ? ch->PG[dny + 1] : 0;
      const double PBy = (nx == 0) ? 1 : pb;
      const int ng2 = ngood - (t->ngv[l] > 0) - (t->ngv[r] > 0) + 1;
      const double Pboty = 1.0 / ng2, PDx = 1 - PBx, Pnogx = 1.0 / nnogs;
      pr = ((1 - PGny) * PBy * Pboty) / (PGny * (1 - PGlx) * (1 - PGrx) * PDx * Pnogx);
      if (ia > 0) {
        int v = t->v[nx], c = t->c[nx], bL, bU;
        node_range(t, nx, v, ch->nc[v], &bL, &bU);
        const double qu = (1.0 / p) / (bU - bL + 1);
        const double qx = informed_rule(ch, t, j, nx, l, r, cp.th, inv_s2, ng2, 0, &v, &c);
        pr *= ((1 - ia) * qu + ia * qx) / qu;
      }
      cp.mode = 2;
      cp.nx = nx;
    }

    /* ---- data pass, :467 */
    data_pass(ch, j - 1, j, &pu, j > 0 && !pre, &cp, 1, 0, 0, &st);

    /* ---- Metropolis-Hastings, :469-518 */
    oprev nu_;
    memset(&nu_, 0, sizeof nu_);
    nu_.nx = cp.nx;
    nu_.v = cp.v;
    nu_.c = cp.c;
    for (int s = 0; s < ORC_NODES; s++) nu_.tho[s] = cp.th[s];
    double S0[ORC_NODES], S1[ORC_NODES];
    memcpy(S0, st.S0, sizeof S0);
    memcpy(S1, st.S1, sizeof S1);
    const double lu = log((double)orc_unif(&ch->rng));
    double lalpha = 0;
    int accepted = 0, capped = 0;
    if (cp.mode == 1) {
      if (st.nL >= 5 && st.nR >= 5) {
        const double bl = st.sL0 * inv_s2, Ml = st.sL1 * inv_s2, br = st.sR0 * inv_s2, Mr = st.sR1 * inv_s2;
        lalpha = log(pr) + lh(bl, Ml, t2) + lh(br, Mr, t2) - lh(bl + br, Ml + Mr, t2);
        if (lalpha > 0) lalpha = 0;
        if (lu < lalpha) {
          const int L = first_free(t->used, -1, ch->node_cap);
          const int R = L < 0 ? -1 : first_free(t->used, L, ch->node_cap);
          if (L < 0 || R < 0) {
            ch->n_cap++;
            capped = 1;
          } else {
            const int nx = cp.nx;
            t->v[nx] = (uint8_t)cp.v;
            t->c[nx] = (uint8_t)cp.c;
            t->l[nx] = (uint8_t)L;
            t->r[nx] = (uint8_t)R;
            t->par[L] = t->par[R] = (uint8_t)nx;
            t->l[L] = t->r[L] = t->l[R] = t->r[R] = 0;
            t->d[L] = t->d[R] = (uint8_t)(t->d[nx] + 1);
            t->used |= (1ull << L) | (1ull << R);
            t->leaf = (t->leaf & ~(1ull << nx)) | (1ull << L) | (1ull << R);
            t->ngv[L] = (uint8_t)count_goodvars(t, L, ch->nc, p);
            t->ngv[R] = (uint8_t)count_goodvars(t, R, ch->nc, p);
            S0[L] = st.sL0;
            S1[L] = st.sL1;
            S0[R] = st.sR0;
            S1[R] = st.sR1;
            nu_.mode = 1;
            nu_.L = L;
            nu_.R = R;
            ch->n_birth++;
            accepted = 1;
          }
        }
      }
    } else if (cp.mode == 2) {
      const int nx = cp.nx, l = t->l[nx], r = t->r[nx];
      const double bl = S0[l] * inv_s2, Ml = S1[l] * inv_s2, br = S0[r] * inv_s2, Mr = S1[r] * inv_s2;
      lalpha = log(pr) + lh(bl + br, Ml + Mr, t2) - lh(bl, Ml, t2) - lh(br, Mr, t2);
      if (lalpha > 0) lalpha = 0;
      if (lu < lalpha) {
        S0[nx] = S0[l] + S0[r];
        S1[nx] = S1[l] + S1[r];
        t->used &= ~((1ull << l) | (1ull << r));
        t->leaf = (t->leaf & ~((1ull << l) | (1ull << r))) | (1ull << nx);
        t->l[nx] = t->r[nx] = t->v[nx] = t->c[nx] = 0;
        nu_.mode = 2;
        nu_.l = l;
        nu_.r = r;
        ch->n_death++;
        accepted = 1;
      }
    }
    /* ---- leaf values: one normal per slot, used leaves only keep theirs, :520-528 (node_cap plays the role of the
     * reference's NODES: with NODES slots it draws NODES normals) */
    for (int s = 0; s < ch->node_cap; s++) {
      const double z = (double)orc_rnorm(&ch->rng);
      if (BIT(t->leaf, s) && BIT(t->used, s)) {
        const double b = S0[s] * inv_s2, M = S1[s] * inv_s2;
        t->th[s] = (float)(M / (ch->a_prior + b) + z / sqrt(ch->a_prior + b));
      }
    }
    for (int s = 0; s < ORC_NODES; s++) nu_.thn[s] = (double)t->th[s];
    if (ch->trace && ch->trace_len < ch->trace_cap) {
      orc_trace* tr = &ch->trace[ch->trace_len++];
      tr->it = it;
      tr->j = j;
      tr->mode = cp.mode;
      tr->nx = cp.nx;
      tr->v = cp.v;
      tr->c = cp.c;
      tr->nl = st.nL;
      tr->nr = st.nR;
      tr->accepted = accepted;
      tr->ngood = ngood;
      tr->nnogs = nnogs;
      tr->cap = capped;
      tr->pr = pr;
      tr->lalpha = lalpha;
      tr->lu = lu;
      tr->sl0 = st.sL0;
      tr->sl1 = st.sL1;
      tr->sr0 = st.sR0;
      tr->sr1 = st.sR1;
      memcpy(tr->S0, st.S0, sizeof tr->S0);
      memcpy(tr->S1, st.S1, sizeof tr->S1);
      memcpy(tr->rng_after, ch->rng.s, sizeof tr->rng_after);
    }
    pu = nu_;
  }
  /* ---- closing pass: apply the last tree, residual sum of squares, posterior-mean sum, :536-541 */
  const int keep = it >= ch->nskip;
  ocur dummy;
  memset(&dummy, 0, sizeof dummy);
  data_pass(ch, m - 1, m, &pu, 1, &dummy, 0, 1, keep, &st);
  if (ch->probit) {
    /* latent redraw given the fit, rows in order, :543-547 */
    for (int i = 0; i < ch->n; i++) {
      const double f = ch->AF[i];
      ch->Y[i] = ch->ybin[i] == 0 ? -orc_rtnorm(&ch->rng, -f, (double)ch->offset)
                                  : orc_rtnorm(&ch->rng, f, -(double)ch->offset);
    }
  } else {
    /* inverse chi-square via Gamma((n+nu)/2): nu*lambda is a float product, n+nu a float sum, :550-557 */
    const float nul = ch->nu * ch->lambda;
    const float nnu = (float)ch->n + ch->nu;
    ch->sigma = sqrt(((double)nul + st.rss) / (2.0 * orc_rgamma(&ch->rng, (double)nnu / 2.0)));
  }
  if (sigma_out) sigma_out[it] = (float)ch->sigma;
  if (keep && keep_trees && trees_out) {
    uint64_t* base = trees_out + (size_t)(it - ch->nskip) * m * ORC_NODES;
    for (int j = 0; j < m; j++)
      for (int s = 0; s < ORC_NODES; s++) base[(size_t)j * ORC_NODES + s] = pack_slot(&ch->T[j], s);
  }
  ch->it++;
}

int orc_chain_sweeps(orc_chain* ch, int k, float* sigma_out, uint64_t* trees_out, int keep_trees) {
  if (!ch) return -1;
  const int total = ch->nskip + ch->ndpost;
  for (int s = 0; s < k && ch->it < total; s++) one_sweep(ch, sigma_out, trees_out, keep_trees);
  return 0;
}

/* trmean = (float)TM / (float)ndpost, :573-578 */
void orc_chain_trmean(const orc_chain* ch, float* trmean_out) {
  for (int i = 0; i < ch->n; i++) trmean_out[i] = (float)ch->TM[i] / (float)ch->ndpost;
}
void orc_chain_stats(const orc_chain* ch, int* st) {
  st[0] = ch->n_birth;
  st[1] = ch->n_death;
  st[2] = ch->n_cap;
}
int orc_chain_iter(const orc_chain* ch) { return ch->it; }
void orc_chain_rng(const orc_chain* ch, uint32_t* s4) { memcpy(s4, ch->rng.s, 16); }
double orc_chain_sigma(const orc_chain* ch) { return ch->sigma; }
void orc_chain_af(const orc_chain* ch, double* af) { memcpy(af, ch->AF, sizeof(double) * ch->n); }
void orc_chain_latent(const orc_chain* ch, double* y) { memcpy(y, ch->Y, sizeof(double) * ch->n); }
void orc_chain_leaf(const orc_chain* ch, uint8_t* leaf) { memcpy(leaf, ch->leaf, (size_t)ch->m * ch->n); }
void orc_chain_pack_trees(const orc_chain* ch, uint64_t* words) {
  for (int j = 0; j < ch->m; j++)
    for (int s = 0; s < ORC_NODES; s++) words[(size_t)j * ORC_NODES + s] = pack_slot(&ch->T[j], s);
}
long orc_chain_trace_len(const orc_chain* ch) { return ch->trace_len; }

int orc_fit(const uint8_t* q, const float* y, const float* wt, const int* numcut, const float* pgtab,
            float* sigma_out, float* trmean_out, uint64_t* trees_out,
            int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
            unsigned seed, int keep_trees, int* stats_out, int probit, const int* ybin, float offset,
            const orc_opts* opts) {
  orc_chain* ch = orc_chain_new(q, y, wt, numcut, pgtab, n, p, m, nskip, ndpost, tau, lambda, nu, sigest, seed, probit,
                                ybin, offset, opts);
  if (!ch) return -1;
  orc_chain_sweeps(ch, nskip + ndpost, sigma_out, trees_out, keep_trees);
  if (trmean_out) orc_chain_trmean(ch, trmean_out);
  if (stats_out) orc_chain_stats(ch, stats_out);
  orc_chain_free(ch);
  return 0;
}

/* ------------------------------------------------------------------ glue restatement (src/glue.cpp:29-57) */
typedef struct sbuf {
  char* p;
  size_t len, cap;
} sbuf;
static void sb_put(sbuf* b, const char* s, size_t k) {
  if (b->len + k < b->cap) memcpy(b->p + b->len, s, k);
  b->len += k;
}
static void sb_printf(sbuf* b, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
#include <stdarg.h>
static void sb_printf(sbuf* b, const char* fmt, ...) {
  char tmp[128];
  va_list ap;
  va_start(ap, fmt);
  const int k = vsnprintf(tmp, sizeof tmp, fmt, ap);
  va_end(ap);
  sb_put(b, tmp, (size_t)k);
}

/* pre-order walk with ids root 1, children 2id / 2id+1; internal nodes print "v c 0", leaves "0 0 theta";
 * doubles as %.10g (an ostream with precision(10)). varcount counts internal nodes per variable. */
size_t orc_tree_text(const uint64_t* trees, int ndpost, int m, int p, char* buf, size_t cap, int* varcount) {
  sbuf b = {buf, 0, cap};
  sb_printf(&b, "%d %d %d\n", ndpost, m, p);
  if (varcount) memset(varcount, 0, sizeof(int) * (size_t)ndpost * p);
  int st_s[ORC_NODES + 2];
  uint64_t st_id[ORC_NODES + 2];
  int ord_s[ORC_NODES];
  uint64_t ord_id[ORC_NODES];
  for (int dr = 0; dr < ndpost; dr++)
    for (int j = 0; j < m; j++) {
      const uint64_t* T = trees + ((size_t)dr * m + j) * ORC_NODES;
      int top = 0, cnt = 0;
      st_s[top] = 0;
      st_id[top++] = 1;
      while (top > 0 && cnt < ORC_NODES) {
        const int s = st_s[--top];
        const uint64_t id = st_id[top];
        ord_s[cnt] = s;
        ord_id[cnt++] = id;
        const int L = (int)((T[s] >> 16) & 255), R = (int)((T[s] >> 24) & 255);
        if (L) {
          st_s[top] = R;
          st_id[top++] = 2 * id + 1;
          st_s[top] = L;
          st_id[top++] = 2 * id;
        }
      }
      sb_printf(&b, "%d\n", cnt);
      for (int k = 0; k < cnt; k++) {
        const int s = ord_s[k];
        const int leaf = ((T[s] >> 16) & 255) == 0;
        const int v = leaf ? 0 : (int)(T[s] & 255), c = leaf ? 0 : (int)((T[s] >> 8) & 255);
        if (!leaf && varcount && v < p) varcount[(size_t)dr * p + v] += 1;
        double th = 0.0;
        if (leaf) {
          const uint32_t u = (uint32_t)(T[s] >> 32);
          float f;
          memcpy(&f, &u, 4);
          th = (double)f;
        }
        sb_printf(&b, "%llu %d %d %.10g\n", (unsigned long long)ord_id[k], v, c, th);
      }
    }
  if (b.len < cap) buf[b.len] = 0;
  return b.len;
}

void orc_yhat(const uint64_t* trees, const uint8_t* q, int n, int ndpost, int m, double* yhat) {
  for (int dr = 0; dr < ndpost; dr++) {
    double* f = yhat + (size_t)dr * n;
    for (int i = 0; i < n; i++) f[i] = 0.0;
    for (int j = 0; j < m; j++) {
      const uint64_t* T = trees + ((size_t)dr * m + j) * ORC_NODES;
      for (int i = 0; i < n; i++) {
        int s = 0;
        while ((T[s] >> 16) & 255) {
          const int v = (int)(T[s] & 255), c = (int)((T[s] >> 8) & 255);
          s = q[(size_t)v * n + i] <= c ? (int)((T[s] >> 16) & 255) : (int)((T[s] >> 24) & 255);
        }
        const uint32_t u = (uint32_t)(T[s] >> 32);
        float th;
        memcpy(&th, &u, 4);
        f[i] += (double)th;
      }
    }
  }
}

/* ------------------------------------------------------------------ predictor restatement (src/predict.cpp) */
typedef struct { uint64_t id; int v, c; double th; } onode;

static int find_id(const onode* nd, int nn, uint64_t id) {
  for (int k = 0; k < nn; k++)
    if (nd[k].id == id) return k;
  return -1;
}

/* text -> flat arrays, breadth first from id 1, the two children of a node stored next to each other;
 * cut = cutpoints[v][c]. cutpoints is the concatenation of the per-variable arrays, cut_off[v] their offsets. */
int orc_compile_trees(const char* text, const double* cutpoints, const int* cut_off, int pcut, orc_ct* out) {
  memset(out, 0, sizeof *out);
  const char* s = text;
  char* e;
  out->ndpost = (int)strtol(s, &e, 10); s = e;
  out->ntree = (int)strtol(s, &e, 10); s = e;
  out->p = (int)strtol(s, &e, 10); s = e;
  const long ntr = (long)out->ndpost * out->ntree;
  if (ntr < 0) return -1;
  size_t cap = 1024, len = 0;
  out->var = (int*)malloc(cap * sizeof(int));
  out->child = (int*)malloc(cap * sizeof(int));
  out->cut = (double*)malloc(cap * sizeof(double));
  out->theta = (double*)malloc(cap * sizeof(double));
  out->root = (int*)malloc((ntr > 0 ? ntr : 1) * sizeof(int));
  onode* nd = NULL;
  int* qk = NULL;
  int* slot = NULL;
  int ndcap = 0;
  for (long t = 0; t < ntr; t++) {
    const int nn = (int)strtol(s, &e, 10);
    if (e == s || nn < 1) goto fail;
    s = e;
    if (nn > ndcap) {
      ndcap = nn;
      nd = (onode*)realloc(nd, sizeof(onode) * nn);
      qk = (int*)realloc(qk, sizeof(int) * nn);
      slot = (int*)realloc(slot, sizeof(int) * nn);
    }
    for (int k = 0; k < nn; k++) {
      nd[k].id = strtoull(s, &e, 10); s = e;
      nd[k].v = (int)strtol(s, &e, 10); s = e;
      nd[k].c = (int)strtol(s, &e, 10); s = e;
      nd[k].th = strtod(s, &e);
      if (e == s) goto fail;
      s = e;
    }
    if (len + 2 * (size_t)nn + 2 > cap) {
      while (len + 2 * (size_t)nn + 2 > cap) cap *= 2;
      out->var = (int*)realloc(out->var, cap * sizeof(int));
      out->child = (int*)realloc(out->child, cap * sizeof(int));
      out->cut = (double*)realloc(out->cut, cap * sizeof(double));
      out->theta = (double*)realloc(out->theta, cap * sizeof(double));
    }
    int qn = 0;
    qk[qn] = find_id(nd, nn, 1);
    if (qk[0] < 0) goto fail;
    slot[qn++] = (int)len;
    out->root[t] = (int)len;
    out->var[len] = 0; out->child[len] = -1; out->cut[len] = 0; out->theta[len] = 0; len++;
    for (int h = 0; h < qn; h++) {
      const int k = qk[h], me = slot[h];
      const int kl = find_id(nd, nn, 2 * nd[k].id);
      out->theta[me] = nd[k].th;
      if (kl >= 0) {
        const int kr = find_id(nd, nn, 2 * nd[k].id + 1);
        if (kr < 0 || qn + 2 > nn || nd[k].v < 0 || nd[k].v >= pcut) goto fail;
        const int ch = (int)len;
        out->var[me] = nd[k].v;
        out->cut[me] = cutpoints[cut_off[nd[k].v] + nd[k].c];
        out->child[me] = ch;
        for (int z = 0; z < 2; z++) {
          out->var[len] = 0; out->child[len] = -1; out->cut[len] = 0; out->theta[len] = 0; len++;
        }
        qk[qn] = kl; slot[qn++] = ch;
        qk[qn] = kr; slot[qn++] = ch + 1;
      }
    }
  }
  out->nnodes = (int)len;
  free(nd); free(qk); free(slot);
  return 0;
fail:
  free(nd); free(qk); free(slot);
  orc_ct_free(out);
  return -2;
}

void orc_ct_free(orc_ct* ct) {
  free(ct->var); free(ct->child); free(ct->cut); free(ct->theta); free(ct->root);
  memset(ct, 0, sizeof *ct);
}

/* tree-outer, row-inner; each row adds the trees in order, then mu; x >= cut goes right. src/predict.cpp:49-73 */
int orc_predict_draw(const orc_ct* ct, const double* const* cols, int n, int p, int m_draw, double mu, double* out) {
  if (p != ct->p) return -1;
  if (m_draw < 1 || m_draw > ct->ndpost) return -2;
  for (int i = 0; i < n; i++) out[i] = 0.0;
  const int* r0 = ct->root + (size_t)(m_draw - 1) * ct->ntree;
  for (int t = 0; t < ct->ntree; t++) {
    const int k0 = r0[t];
    for (int i = 0; i < n; i++) {
      int k = k0;
      while (ct->child[k] >= 0) k = ct->child[k] + (int)(cols[ct->var[k]][i] >= ct->cut[k]);
      out[i] += ct->theta[k];
    }
  }
  for (int i = 0; i < n; i++) out[i] += mu;
  return 0;
}

/* ------------------------------------------------------------------ libm images (checks of the device tables) */
/* entries [k0, k0+count) on the generator grid u_k = (k + 1/2) 2^-24 with the reference's own expressions
 * (src/bart_kernel.cpp:59-60, :289, :482): kind 0 logf(u), 1 cosf(6.28318531f*u), 2 log((double)u),
 * 3 sqrtf(-2.f*logf(u)) (float kinds into outf, double kind into outd) */
void orc_libm_image(int kind, long k0, long count, float* outf, double* outd) {
  for (long i = 0; i < count; i++) {
    const uint32_t k = (uint32_t)(k0 + i);
    const float u = ((float)k + 0.5f) * (1.0f / 16777216.0f);
    if (kind == 0) outf[i] = logf(u);
    if (kind == 1) outf[i] = cosf(6.28318531f * u);
    if (kind == 2) outd[i] = log((double)u);
    if (kind == 3) outf[i] = sqrtf(-2.f * logf(u));
  }
}
/* the normal of :60 from two raw generator outputs */
float orc_normal_from(uint32_t x1, uint32_t x2) {
  const float u1 = ((float)(x1 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((float)(x2 >> 8) + 0.5f) * (1.0f / 16777216.0f);
  return sqrtf(-2.f * logf(u1)) * cosf(6.28318531f * u2);
}
