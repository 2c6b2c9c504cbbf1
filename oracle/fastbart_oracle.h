/* fastbart_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C, of the reference BART sampler (src/bart_kernel.cpp, FAST_CPU build, internal
 * xoshiro128** generator = FASTBART_RNG=internal), of the output glue (src/glue.cpp:29-57) and of the predictor
 * (src/predict.cpp). It is the checker the tests, the smoke test and the CPU leg of the benchmark compare the HIP
 * library against; nothing in the product links or calls it.
 *
 * Differences in structure (not in results) from the reference: chain state lives in a heap struct (reentrant, many
 * chains per process), rows are stored in natural order and the 16 summation lanes are kept as lane = row % 16, and a
 * second statistics mode (FX64, int64 fixed point) defines the order-independent variant used for large chains.
 *
 * Pinning (how each part is checked, see tests/): golden trajectories written by the compiled reference core
 * (oracle/_ref), KATs for the generator, exhaustive comparison of the normal draw against the host libm, brute force for
 * the tree helpers and the predictor.
 */
#ifndef FASTBART_ORACLE_H
#define FASTBART_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NODES 64     /* node slots per tree, src/bart_kernel.h:22-24 */
#define ORC_KL 16        /* summation lanes, src/bart_kernel.h:25-27 */
#define ORC_MAXDEPTH 24  /* src/bart_kernel.h:33 */
#define ORC_PMAX 32      /* src/bart_kernel.h:20 */
#define ORC_MMAX 200     /* src/bart_kernel.h:21 */
#define ORC_NMAX 65536   /* src/bart_kernel.h:15-16 */

enum { ORC_STATS_REF16 = 0, ORC_STATS_FX64 = 1 };
/* FX64 fixed-point scales: w and w*r at 2^32, the residual sum of squares at 2^24 */
#define ORC_FX_SHIFT 32
#define ORC_FX_RSS_SHIFT 24

typedef struct orc_rng { uint32_t s[4]; } orc_rng;

void orc_rng_seed(orc_rng* r, uint32_t seed);      /* src/bart_kernel.cpp:381-382 */
uint32_t orc_rng_next(orc_rng* r);                 /* :53-58 */
float orc_unif(orc_rng* r);                        /* :59 */
float orc_rnorm(orc_rng* r);                       /* :60 */
double orc_rgamma(orc_rng* r, double a);           /* :73-83 */
double orc_rtnorm(orc_rng* r, double mean, double tau); /* :282-292 */
/* log of the libm comparisons of rtnorm and rgamma (see fastbart_oracle.c); buf holds cap records of 7 doubles */
void orc_cmp_log(double* buf, long cap, long sample);
long orc_cmp_count(long* seen);

/* one record per tree update, filled when a trace buffer is given (debugging / kernel-level parity) */
typedef struct orc_trace {
  int32_t it, j, mode, nx, v, c, nl, nr, accepted, ngood, nnogs, cap;
  double pr, lalpha, lu;
  double sl0, sl1, sr0, sr1;
  double S0[ORC_NODES], S1[ORC_NODES];
  uint32_t rng_after[4];
} orc_trace;

typedef struct orc_opts {
  int stats_mode;     /* ORC_STATS_REF16 (default) or ORC_STATS_FX64 */
  int node_cap;       /* node slots per tree, 3..64 (0 = 64): the reference's compile-time NODES
                       * (src/bart_kernel.h:22-24), i.e. slots a tree may use and normals drawn per update */
  orc_trace* trace;   /* optional, capacity trace_cap records */
  long trace_cap;
} orc_opts;

typedef struct orc_chain orc_chain;

/* Create a chain with the same inputs as bart_fit_cpu (src/bart_kernel.h:63-66). Buffers are copied. */
orc_chain* orc_chain_new(const uint8_t* q, const float* y, const float* wt, const int* numcut, const float* pgtab,
                         int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
                         unsigned seed, int probit, const int* ybin, float offset, const orc_opts* opts);
void orc_chain_free(orc_chain* ch);
/* run sweeps [it, it+k); outputs are written for the sweeps covered (arrays sized as for bart_fit_cpu) */
int orc_chain_sweeps(orc_chain* ch, int k, float* sigma_out, uint64_t* trees_out, int keep_trees);
void orc_chain_trmean(const orc_chain* ch, float* trmean_out);
void orc_chain_stats(const orc_chain* ch, int* stats_out);
int orc_chain_iter(const orc_chain* ch);
/* state accessors for parity checks */
void orc_chain_rng(const orc_chain* ch, uint32_t* s4);
double orc_chain_sigma(const orc_chain* ch);
void orc_chain_af(const orc_chain* ch, double* af);            /* [n] running fit */
void orc_chain_latent(const orc_chain* ch, double* y);         /* [n] current y (latent z for probit) */
void orc_chain_leaf(const orc_chain* ch, uint8_t* leaf);       /* [m][n] leaf slot of each row per tree */
void orc_chain_pack_trees(const orc_chain* ch, uint64_t* words); /* [m][64] current trees as output words */
long orc_chain_trace_len(const orc_chain* ch);
/* proposal study (no reference counterpart): births draw their rule with probability alpha from the rule's
 * conditional given the tree, sigma and the partial residuals, the mixture's probability entering the ratio of births
 * and deaths; 0 restores the reference's proposal */
void orc_chain_set_informed(orc_chain* ch, double alpha);
long orc_chain_informed_births(const orc_chain* ch);
/* that conditional at one node given directly (no reference counterpart): rows with member[i] != 0, weights w,
 * partial residuals r, cutpoint ranks q [p][n]; ranges[v] = L | U << 8 (U < L: no cut); the node's depth and
 * good-variable count, the tree's good-leaf count before the move, the depth prior pg [64], sigma, tau.
 * q_out [p][255] = q(v, c), *entropy = its entropy (-1 if no rule qualifies); returns the number of rules */
int orc_split_conditional(int n, int p, const uint8_t* q, const uint8_t* member, const double* w, const double* r,
                          const uint16_t* ranges, int depth, int ngv, int ngood, const double* pg, double sigma,
                          double tau, double* q_out, double* entropy);

/* whole fit, same contract as bart_fit_cpu with the internal generator; returns 0 or a negative error */
int orc_fit(const uint8_t* q, const float* y, const float* wt, const int* numcut, const float* pgtab,
            float* sigma_out, float* trmean_out, uint64_t* trees_out,
            int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
            unsigned seed, int keep_trees, int* stats_out, int probit, const int* ybin, float offset,
            const orc_opts* opts);

/* ---- glue (src/glue.cpp:29-57) ---- */
/* treedraws text; returns the length (without NUL); writes at most cap bytes (NUL-terminated when it fits).
 * varcount [ndpost][p] is filled when non-null. */
size_t orc_tree_text(const uint64_t* trees, int ndpost, int m, int p, char* buf, size_t cap, int* varcount);
/* yhat[dr][i] = sum over trees in order of the leaf value reached by row i (q <= c goes left) */
void orc_yhat(const uint64_t* trees, const uint8_t* q, int n, int ndpost, int m, double* yhat);

/* ---- predictor (src/predict.cpp) ---- */
typedef struct orc_ct {
  int ndpost, ntree, p, nnodes;
  int* var; int* child; double* cut; double* theta; int* root;
} orc_ct;
int orc_compile_trees(const char* text, const double* cutpoints, const int* cut_off, int pcut, orc_ct* out);
void orc_ct_free(orc_ct* ct);
/* cols: p column pointers; m_draw 1-based; out [n] */
int orc_predict_draw(const orc_ct* ct, const double* const* cols, int n, int p, int m_draw, double mu, double* out);

/* ---- libm images on the generator grid (checks of the device tables) ---- */
void orc_libm_image(int kind, long k0, long count, float* outf, double* outd);
float orc_normal_from(uint32_t x1, uint32_t x2);

#ifdef __cplusplus
}
#endif
#endif
