/* fastbart_hip.h -- C-ABI of the MI355X (gfx950) BART sampler library libfastbart_hip.so.
 *
 * Plain C types and caller-owned host buffers only. Every entry point returns an int status (0 = success,
 * negative = error; fbh_last_error() describes the last error of the calling thread) except the three drop-in
 * symbols, which keep the reference's void signatures.
 *
 * Drop-in boundary (reference interfaces this library replaces):
 *   bart_fit_cpu / bart_wfit / bart_set_rng   src/bart_kernel.h:49-70 (called from src/glue.cpp:24-27)
 *   fbh_fit_batch / fbh_batch_*               new: many independent chains in one persistent launch (no reference
 *                                             counterpart; the reference runs one chain per R process)
 *   fbh_tree_text / fbh_yhat                  the output half of fbart_cpp, src/glue.cpp:29-57
 *   fbh_compile_trees / fbh_predict_draw(s)   compile_trees_cpp / predict_draw_cpp, src/predict.cpp:12-73
 */
#ifndef FASTBART_HIP_H
#define FASTBART_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBH_NODES 64       /* node slots per tree (src/bart_kernel.h:22-24) */
#define FBH_NMAX 65536     /* rows per chain (src/bart_kernel.h:15-16) */
#define FBH_PMAX 32        /* predictors (src/bart_kernel.h:20) */
#define FBH_MMAX 200       /* trees (src/bart_kernel.h:21) */

/* error codes */
#define FBH_OK 0
#define FBH_EINVAL -1      /* bad argument / shape outside the limits above */
#define FBH_ENODEV -2      /* no usable gfx950 device */
#define FBH_EHIP -3        /* a HIP call failed */
#define FBH_ENOMEM -4      /* device or host allocation failed */
#define FBH_ERNG -5        /* an external (R) generator is installed: the device cannot reproduce its stream */
#define FBH_EPARSE -6      /* malformed tree text */

/* statistics modes */
#define FBH_STATS_REF16 0  /* double sums in 16 row lanes, bit-identical to the reference CPU sampler */
#define FBH_STATS_FX64 1   /* int64 fixed point (w, w*r at 2^32; rss at 2^24): order independent. Range: every chain
                              needs n max(wt) 4 max|y| < 2^30 and n max(wt) (4 max|y|)^2 < 2^38 (probit: 64 in place
                              of 4 max|y|), so that the int64 sums cannot wrap; fbh_batch_create returns FBH_EINVAL
                              otherwise (rescale y, or use REF16) */

/* ---------------------------------------------------------------- drop-in single chain */
/* Same arguments and semantics as src/bart_kernel.h:63-66 / src/bart_kernel.cpp:304-307, run on the GPU with the
 * internal xoshiro128** generator. If bart_set_rng installed an external generator, nothing is computed,
 * stats_out[0..2] are set to -1 and fbh_last_status() returns FBH_ERNG. */
void bart_fit_cpu(const uint8_t* qin, const float* yin, const float* wtin, const int* numcut, const float* pgtab,
                  float* sigma_out, float* trmean_out, uint64_t* trees_out,
                  int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
                  unsigned seed, int keep_trees, int* stats_out, int probit, const int* ybin, float offset);
/* src/bart_kernel.h:49-58: bart_fit_cpu with probit = 0 */
void bart_wfit(const uint8_t* qin, const float* yin, const float* wtin, const int* numcut, const float* pgtab,
               float* sigma_out, float* trmean_out, uint64_t* trees_out,
               int n, int p, int m, int nskip, int ndpost, float tau, float lambda, float nu, float sigest,
               unsigned seed, int keep_trees, int* stats_out);
/* src/bart_kernel.h:69: all-null selects the internal generator (FASTBART_RNG=internal); any non-null pointer marks
 * an external generator, which the device path refuses (FBH_ERNG). */
void bart_set_rng(double (*u)(void), double (*nrm)(void), double (*e)(void), double (*chisq)(double));
/* status of the last drop-in call on this thread */
int fbh_last_status(void);
const char* fbh_last_error(void);

/* ---------------------------------------------------------------- batched chains */
typedef struct fbh_chain_in {
  int n, p;                 /* rows, predictors */
  const uint8_t* q;         /* [p][n] cutpoint ranks: x < xi[v][c] <=> q <= c */
  const float* y;           /* [n] centred outcome (ignored for probit) */
  const float* wt;          /* [n] 1/w^2 */
  const int* numcut;        /* [p] */
  const int* ybin;          /* [n] 0/1 outcomes, probit only */
  float tau, lambda, nu, sigest, offset;
  unsigned seed;            /* xoshiro128** seed of this chain */
} fbh_chain_in;

typedef struct fbh_opts {
  int m, nskip, ndpost;     /* trees, burn-in sweeps, kept sweeps (shared by the batch) */
  int probit;               /* 0 wbart, 1 pbart */
  int keep_trees;           /* store the kept trees on the device */
  const float* pgtab;       /* [64] base/(1+d)^power as float */
  int stats_mode;           /* FBH_STATS_REF16 / FBH_STATS_FX64 */
  int node_cap;             /* node slots a tree may use (3..64), 0 = 64 */
  int device;               /* HIP device ordinal */
  int sweeps_per_launch;    /* 0 = automatic */
  int block_threads;        /* 0 or 320: one control wave + four data waves per chain (the only shape built) */
  int tile_rows;            /* 0 = automatic; 1024 or 1536 rows staged in LDS per tile (1536 needs the running fit
                               resident in LDS; 1024 alone, with resident = 0, streams it) */
  int row_split;            /* 0 or 1 = one workgroup per chain. W > 1 (FX64 statistics): the rows of every chain
                               are split over W workgroups that run the same control steps and add their fixed-point
                               pass totals after every pass -- exact, as fixed-point sums are order-free, so the
                               results equal the one-workgroup FX64 chain bit for bit. Every chain needs n >= 64 W,
                               and all C x W workgroups must fit on the device at once (FBH_EINVAL if not). Probit
                               chains split too, their members holding multiples of 64 rows (every member non-empty)
                               and their latents redrawn by the device-wide kernels (probit_redraw 2; 1 is refused).
                               -1 = automatic for FX64 batches: the largest W with C x W <= the CU count and at least
                               3,072 rows per member (1 otherwise). */
  int packing;              /* workgroup -> chain assignment of the sweep kernel: 0 = batch order, 1 = serpentine by
                               rows (large and small chains share CUs), 2 = largest first. Timing only: results do not
                               depend on it. fbh_batch_autotune measures and picks. */
  int graph;                /* 1: fbh_batch_run captures each call's launches into a hipGraph and launches the graph
                               (0: plain stream launches). Results do not depend on it. */
  int chains_per_cu;        /* build of the sweep kernel: 1 = 168 VGPRs, one workgroup per CU; 2 = 128 VGPRs, two
                               workgroups share a CU; 0 = automatic (2 when a wbart batch has more workgroups than
                               the device has CUs). Results do not depend on it. */
  int resident;             /* the running fit during a launch: 1 = streamed from HBM, 2 = resident in LDS (needs the
                               longest chain's 8 n bytes beside the tile within 80 KiB), 0 = automatic (in LDS when
                               it fits: 1024-row tiles for the two-per-CU build, 1536 otherwise) */
  int probit_redraw;        /* probit batches, the latent redraw after every sweep: 1 = inside the sweep kernel (each
                               chain's workgroup walks its rows in windows of 64), 2 = by kernels spread over the
                               device (bands of precomputed decisions, one wave per chain walking them; a launch is
                               then one sweep), 0 = automatic (2 for batches of at most 16 chains). Results do not
                               depend on it. */
} fbh_opts;

typedef struct fbh_chain_out {
  float* sigma;             /* [nskip+ndpost] (may be null) */
  float* trmean;            /* [n] (may be null) */
  uint64_t* trees;          /* [ndpost][m][64] output words (may be null) */
  int* stats;               /* [3] births, deaths, node-cap hits (may be null) */
} fbh_chain_out;

typedef struct fbh_batch fbh_batch;

/* Upload the inputs and initialise every chain (root-only trees, seeded generator, probit latents). */
int fbh_batch_create(const fbh_chain_in* chains, int nchains, const fbh_opts* opts, fbh_batch** out);
/* Enqueue the next nsweeps sweeps of every chain on the batch's stream (asynchronous). */
int fbh_batch_run(fbh_batch* b, int nsweeps);
/* Wait for the batch's work; FBH_EHIP if a row-split exchange timed out (its results are invalid). */
int fbh_batch_sync(fbh_batch* b);
/* sweeps done so far */
int fbh_batch_iter(const fbh_batch* b);
/* workgroups per chain (the row split in effect; 1 without one) */
int fbh_batch_row_split(const fbh_batch* b);

/* ---------------------------------------------------------------- launch configuration and autotuner (new; the
 * reference runs one chain per process and has no launch to configure) */
#define FBH_TUNE_MAX 12
typedef struct fbh_tune {
  int packing, tile_rows, resident, chains_per_cu;  /* the configuration chosen (in effect on return) */
  int ncand;                                        /* candidates timed (both stages) */
  int cand_packing[FBH_TUNE_MAX], cand_tile_rows[FBH_TUNE_MAX], cand_resident[FBH_TUNE_MAX],
      cand_chains_per_cu[FBH_TUNE_MAX];
  double ms_per_sweep[FBH_TUNE_MAX];
  int sweeps;                                       /* sweeps the tuning ran: real sweeps, results unchanged */
  int from_cache;                                   /* 1: taken from the autotuner cache, nothing timed */
} fbh_tune;
/* Switch packing (0, 1, 2 as fbh_opts.packing), tile rows and resident (as in fbh_opts; 0 = the automatic choice)
 * and the build (chains_per_cu 1 or 2, 0 = keep) between runs; waits for the batch's queued work. Results do not
 * depend on any of them. */
int fbh_batch_configure(fbh_batch* b, int packing, int tile_rows, int resident, int chains_per_cu);
/* A packing of the caller's: logical chain order[w] runs on workgroup w (on workgroups w W .. w W + W - 1 with a row
 * split of W); order is a permutation of 0..nchains-1. Waits for the batch's queued work; results do not depend on it. */
int fbh_batch_set_order(fbh_batch* b, const int* order);
/* Probit batches: the latent redraw (mode 1, 2, 0 = automatic, as fbh_opts.probit_redraw) from the next run on;
 * slots = 0 for the default stream buffer of mode 2 (twice the longest chain's rows and its widest band), or a
 * smaller multiple of 64: slots past it are then generated by the walk itself (slower; the results are the same).
 * Waits for the batch's queued work. FBH_EINVAL for a wbart batch. */
int fbh_batch_probit_redraw(fbh_batch* b, int mode, int slots);
/* The redraw in effect (mode 0 for a wbart batch), the mode-2 stream buffer's slots per chain (0 before its first
 * use), and the attempts mode 2 has decided outside the precomputed bands so far, summed over chains (synchronises). */
int fbh_batch_probit_info(fbh_batch* b, int* mode, int* slots, unsigned long long* spot);
/* the configuration in effect (packing -1 after fbh_batch_set_order; resident 1 streamed, 2 in LDS) */
int fbh_batch_config(const fbh_batch* b, int* packing, int* tile_rows, int* resident, int* chains_per_cu, int* graph);
/* Time launch configurations on the batch's next sweeps and keep the fastest: first build x tile kind (1024-row tiles
 * streamed or in LDS, 1536-row tiles in LDS) at largest-first packing, then the three packings with the winner; one
 * warm-up launch, then two launches of trial_sweeps sweeps per candidate. The sweeps are real sweeps of every chain
 * and count toward nskip + ndpost (FBH_EINVAL if too few are left); no candidate changes any result. out may be
 * null. */
int fbh_batch_autotune(fbh_batch* b, int trial_sweeps, fbh_tune* out);
/* fbh_batch_autotune with a cache file (a tile schedule per batch shape): if cache_path holds an entry for this
 * batch's shape -- device architecture and CU count, library build, statistics mode, probit, row split, chain count,
 * trees, node cap, rows per chain (minimum, quartiles, maximum in 256-row steps), predictors per chain (range) -- its
 * configuration is applied at once (no sweep run, out->from_cache = 1, out->ncand = 0); otherwise the batch is tuned
 * as by fbh_batch_autotune and one line is appended to the file (created if absent). Lines are
 * "<key>\t<packing> <tile_rows> <resident> <chains_per_cu> <ms per sweep>"; the last line for a key wins; an entry
 * this batch cannot run is tuned afresh. A null or empty cache_path is fbh_batch_autotune. FBH_EINVAL if the file
 * cannot be opened for appending (before any sweep) or written (the tuned configuration is then in effect). */
int fbh_batch_autotune_cached(fbh_batch* b, int trial_sweeps, const char* cache_path, fbh_tune* out);
/* Copy chain c's outputs to host buffers (synchronises). */
int fbh_batch_fetch(fbh_batch* b, int chain, const fbh_chain_out* out);
/* Copy chain c's sampler state (parity checks, checkpoints): rng[4], sigma, af[n], latent[n], leaf[m][n],
 * tree words [m][64]; any pointer may be null. */
int fbh_batch_state(fbh_batch* b, int chain, uint32_t* rng, double* sigma, double* af, double* latent,
                    uint8_t* leaf, uint64_t* words);
/* yhat[dr][i] = sum over the m trees of draw dr, in tree order, of the leaf value row i reaches (src/glue.cpp:52-56),
 * computed on the device from the stored trees; yhat is [ndpost][n] host memory. */
int fbh_batch_yhat(fbh_batch* b, int chain, double* yhat);
/* Average duration (ms) of the sweep-kernel launches so far, from HIP events on the batch stream, and the launch count.
 * Algorithmic bytes of those launches (37 bytes per row per tree update, (m+1)/m for the closing pass). */
int fbh_batch_timing(const fbh_batch* b, double* avg_ms, int* nlaunch, double* bytes_per_launch);
void fbh_batch_destroy(fbh_batch* b);

/* create + run all sweeps + fetch every chain + destroy */
int fbh_fit_batch(const fbh_chain_in* chains, int nchains, const fbh_opts* opts, const fbh_chain_out* outs);

/* ---------------------------------------------------------------- preprocessing of the R wrappers
 * fwbart's R/fwbart.R:8-27 and fpbart's R/fpbart.R:7-21 in C++, so that callers without R can pass raw x, y, w:
 * BART::bartModelMatrix's default model matrix (constant columns dropped; per column the midpoints of consecutive
 * distinct values when it has at most numcut + 1 of them, else numcut equally spaced points strictly inside
 * [min, max]; xinfo replaces them), q = findInterval(x, xi), fmean, sigest (residual standard error of lm(yc ~ X) when
 * p < n, else sd(yc)), lambda (qchisq), tau, the prior table, the probit offset (qnorm(mean(y))). The results fill an
 * fbh_chain_in (q, y, wt, numcut, ybin, tau, lambda, nu, sigest, offset) and fbh_opts.pgtab. Settings for which the
 * R wrappers fall back to BART (n > 65536, more than 32 predictors after rm.const, ntree > 200, more than 255
 * cutpoints in a column) return FBH_EINVAL. */
typedef struct fbh_prep_opts {
  int ntree, numcut;        /* R's defaults: 200 (fwbart) or 50 (fpbart), 100 */
  int rm_const;             /* 1: drop constant columns */
  double k, power, base;    /* 2, 2, 0.95 */
  double sigdf, sigquant;   /* fwbart: 3, 0.90 */
  double sigest, sigmaf, lambda, fmean;  /* fwbart: NaN = R's NA (derived as R/fwbart.R:21-28 does) */
  double binary_offset;     /* fpbart: NaN = qnorm(mean(y)) */
} fbh_prep_opts;

typedef struct fbh_inputs {  /* owned by the library: fbh_inputs_free */
  int n, p;                 /* rows, predictors kept */
  uint8_t* q;               /* [p][n] cutpoint ranks */
  int* numcut;              /* [p] */
  int* rm_const;            /* [p] kept columns, 1-based (R's rm.const) */
  double* cutpoints;        /* kept columns' cutpoints, concatenated; column v's at cut_off[v] .. cut_off[v+1]-1 */
  int* cut_off;             /* [p+1] */
  float* y;                 /* [n] (float)(y - fmean) for fwbart, zeros for fpbart */
  float* wt;                /* [n] (float)(1 / w^2) for fwbart, ones for fpbart */
  int* ybin;                /* [n] fpbart outcomes (null for fwbart) */
  float pgtab[64];          /* (float)(base / (1 + d)^power) */
  float tau, lambda, nu, sigest, offset;            /* as the sampler takes them */
  double tau_d, lambda_d, nu_d, sigest_d, offset_d; /* the same before the float conversion */
  double fmean;             /* fwbart's mu (0 for fpbart) */
  int lm_rank;              /* rank of [1, X] in the lm fit (0 if sigest was not estimated by lm) */
  void* priv_;
} fbh_inputs;

/* R's defaults for fwbart (probit = 0) or fpbart (probit = 1) */
void fbh_prep_defaults(int probit, fbh_prep_opts* o);
/* x: ncol host columns of n rows; w may be null (all 1); xinfo (may be null): cutpoints of every input column,
 * concatenated, column v's at xinfo_off[v] .. xinfo_off[v+1]-1, each nondecreasing */
int fbh_wbart_inputs(const double* const* x, int n, int ncol, const double* y, const double* w, const double* xinfo,
                     const int* xinfo_off, const fbh_prep_opts* o, fbh_inputs** out);
/* y01: 0/1 outcomes */
int fbh_pbart_inputs(const double* const* x, int n, int ncol, const int* y01, const double* xinfo,
                     const int* xinfo_off, const fbh_prep_opts* o, fbh_inputs** out);
void fbh_inputs_free(fbh_inputs* in);
/* the quantile functions the preprocessing uses: qnorm (AS 241) and qchisq (inverse regularised incomplete gamma) */
double fbh_qnorm(double p);
double fbh_qchisq(double p, double df);

/* ---------------------------------------------------------------- output formats (src/glue.cpp:29-57) */
/* treedraws text of ndpost x m packed trees: "ndpost m p" then per tree the node count and pre-order lines
 * "id v c theta" (%.10g). Returns the full length; writes at most cap bytes (NUL-terminated if it fits).
 * varcount [ndpost][p] may be null. */
size_t fbh_tree_text(const uint64_t* trees, int ndpost, int m, int p, char* buf, size_t cap, int* varcount);
/* yhat[dr][i] from host tree words and q [p][n], evaluated on the device */
int fbh_yhat(const uint64_t* trees, const uint8_t* q, int n, int p, int ndpost, int m, int device, double* yhat);

/* ---------------------------------------------------------------- prediction (src/predict.cpp) */
typedef struct fbh_trees fbh_trees;
/* Parse treedraws text into flat breadth-first child-pair arrays (children adjacent, cut = cutpoints[v][c]) and upload
 * them. cutpoints is the concatenation of the per-variable cutpoint arrays, cut_off[v] the start of variable v
 * (cut_off has pcut+1 entries). */
int fbh_compile_trees(const char* text, size_t len, const double* cutpoints, const int* cut_off, int pcut, int device,
                      fbh_trees** out);
/* ndpost, ntree, p, node count; and a host copy of the flat arrays (any pointer may be null) */
int fbh_trees_info(const fbh_trees* t, int* ndpost, int* ntree, int* p, int* nnodes);
int fbh_trees_arrays(const fbh_trees* t, int* var, int* child, double* cut, double* theta, int* root);
void fbh_trees_free(fbh_trees* t);
/* One draw (m_draw 1-based) for n rows given as p host columns: out[i] = mu + sum over trees in order of theta at the
 * leaf reached with x >= cut going right (predict_draw_cpp, src/predict.cpp:49-73). */
int fbh_predict_draw(const fbh_trees* t, const double* const* cols, int n, int p, int m_draw, double mu, double* out);
/* Batched pwbart: draws[k] (1-based) for k < ndraws over the same rows; out [ndraws][n]. x is row data uploaded
 * once per call as p columns. */
int fbh_predict_draws(const fbh_trees* t, const double* const* cols, int n, int p, const int* draws, int ndraws,
                      double mu, double* out);

/* device-resident rows for repeated prediction: upload once, predict many draws */
typedef struct fbh_rows fbh_rows;
int fbh_rows_upload(const double* const* cols, int n, int p, int device, fbh_rows** out);
void fbh_rows_free(fbh_rows* r);
/* draws (host, 1-based) for the uploaded rows; results copied to out_host [ndraws][n] (predict_draw_cpp,
 * src/predict.cpp:49-73, over many draws) */
int fbh_predict_rows(const fbh_trees* t, const fbh_rows* r, const int* draws, int ndraws, double mu, double* out_host);
/* The same with the draw list and the output in device memory (draws_dev [ndraws], out_dev [ndraws][n]), enqueued on
 * stream (a hipStream_t; null = the default stream) without waiting. A draw outside 1..ndpost yields NaN for its row of
 * out_dev. The handle keeps the ranks of the last rows it predicted in a buffer of its own; calls on different
 * streams are ordered against each other's use of that buffer with events (the output buffers are the caller's). */
int fbh_predict_rows_device(const fbh_trees* t, const fbh_rows* r, const int* draws_dev, int ndraws, double mu,
                            double* out_dev, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): fbh_predict_rows_device without any allocation. The rows' cutpoint
 * ranks go to a fixed per-GPU scratch array of 8 Mi ranks (16 MiB, held by every GPU the library runs on until the
 * process ends) instead of the handle's buffer; r->n x r->p above 8,388,608 is refused (FBH_EINVAL). The trees must be
 * on the device (fbh_compile_trees uploads them when the device is visible). Calls that overlap on one GPU share that
 * scratch and race: order them on one stream. Same outputs as fbh_predict_rows_device, bit for bit. */
int fbh_predict_rows_device_noalloc(const fbh_trees* t, const fbh_rows* r, const int* draws_dev, int ndraws,
                                    double mu, double* out_dev, void* stream);
/* average predict-kernel duration (ms) and launch count since the trees were compiled */
int fbh_predict_timing(const fbh_trees* t, double* avg_ms, int* nlaunch);
/* the kernel the last prediction on t launched: form 1 = rank form (k_predict_rank), 0 = general form (k_predict),
 * -1 = none yet; rows_per_thread = rows each thread of the rank-form kernel evaluated (1 or 2) */
int fbh_predict_last_launch(const fbh_trees* t, int* form, int* rows_per_thread);
/* 1 when the trees and their cutpoint table take the rank-form kernel (cutpoints of every variable nondecreasing and
 * free of NaN, at most 65,535 per variable, trees at most 255 levels deep), 0 when the general kernel is used */
int fbh_trees_rank_form(const fbh_trees* t);
/* device memory for the device-side entry points: allocate, free, copy (kind 0 host->device, 1 device->host,
 * 2 device->device), wait for the device */
int fbh_device_alloc(size_t bytes, int device, void** out);
void fbh_device_free(void* p, int device);
int fbh_device_copy(void* dst, const void* src, size_t bytes, int kind, int device);
int fbh_device_sync(int device);

/* ---------------------------------------------------------------- learned birth proposal: the network
 * No reference counterpart (the reference draws the birth rule uniformly, src/bart_kernel.cpp:433-437). A 3-layer MLP
 * d_in -> h1 -> h2 -> d_out (ReLU), BF16 weights and activations, FP32 accumulation on the matrix cores, FP32 master
 * weights, Adam. Parameters are packed W1 [d_in][h1], b1, W2 [h1][h2], b2, W3 [h2][d_out], b3 (fp32). Batches are host
 * arrays x [M][d_in]; target [M] class indices; mask [M][d_out] (1 = the class may be chosen; null = all). */
typedef struct fbh_mlp fbh_mlp;
/* He-normal weights from seed, zero biases; max_batch bounds M in the calls below */
int fbh_mlp_create(const int* dims, int max_batch, unsigned seed, int device, fbh_mlp** out);
long fbh_mlp_nparams(const fbh_mlp* n);
/* copy the parameters out (get) and/or in (set; resets the Adam state) */
int fbh_mlp_params(fbh_mlp* n, float* get, const float* set);
/* logits [M][d_out] */
int fbh_mlp_forward(fbh_mlp* n, const float* x, int M, float* logits);
/* mean masked-softmax cross-entropy over the batch and its gradient (packed like the parameters; either may be null) */
int fbh_mlp_grad(fbh_mlp* n, const float* x, const int* target, const uint8_t* mask, int M, float* loss, float* grads);
/* steps Adam steps (learning rate lr) on minibatches of batch rows drawn from the M-row dataset with a stream seeded by
 * seed; loss_out [steps] (may be null) receives each step's mean loss */
int fbh_mlp_train(fbh_mlp* n, const float* x, const int* target, const uint8_t* mask, int M, int steps, int batch,
                  float lr, unsigned seed, float* loss_out);
/* average duration (ms) of the device forward passes of fbh_mlp_forward so far, and their count */
int fbh_mlp_timing(const fbh_mlp* n, double* fwd_ms, int* nfwd);
void fbh_mlp_destroy(fbh_mlp* n);

/* ---------------------------------------------------------------- learned birth proposal: sampling with it
 * No reference counterpart. The batch's next sweeps run tree update by tree update over all chains in lock-step:
 * the birth rule comes from the mixture (1 - alpha) x BART's uniform rule + alpha x the network's masked softmax given
 * the node's features (per variable, the split's log-likelihood gain at 16 cut positions from the node's histograms
 * of w and w r over its rows, and the depth), and the Metropolis-Hastings ratios carry the exact proposal
 * correction (births x q_unif / q_mix, deaths x q_mix / q_unif), so the posterior is unchanged. alpha = 0 is BART's
 * proposal. wbart batches only; the network maps 513 features to 8160 = 32 x 255 classes (v * 255 + c). */
/* attach the network (may be null when alpha = 0 and nothing is collected) and the mixture weight; pair_capacity > 0
 * records each accepted birth's (features, allowed rules, rule) for training (the count restarts when it changes) */
int fbh_batch_learned_setup(fbh_batch* b, fbh_mlp* net, float alpha, long pair_capacity);
/* the next nsweeps sweeps with the learned proposal (asynchronous; outputs as for fbh_batch_run) */
int fbh_batch_run_learned(fbh_batch* b, int nsweeps);
/* training pairs collected so far (at most the capacity) */
long fbh_batch_learned_pairs(fbh_batch* b);
/* what a collected pair is: kind 0 (the default) = each accepted birth with BART's rule; kind 1 = the node of a
 * proposed birth (each taken with probability fraction) with the exact split conditional there as a soft target
 * (BART's rule probability times the rule-dependent factors of the birth ratio, over the rules whose children hold at
 * least 5 rows; 32 KB per pair) and a rule drawn from it. Its draws come from a stream seeded by seed apart from the
 * chains' generators, so the chains are unchanged. Set before collecting. */
int fbh_batch_learned_target(fbh_batch* b, int kind, float fraction, unsigned long long seed);
/* kind 1 only (else NaN): the mean over the collected pairs of the conditional's entropy (the cross-entropy floor a
 * network can reach) and of the log of the allowed class count (the uniform rule's cross-entropy), in nats */
int fbh_batch_learned_target_stats(fbh_batch* b, double* mean_entropy, double* mean_log_allowed);
/* keep the last count collected pairs out of fbh_batch_learned_train (for fbh_batch_learned_eval) */
int fbh_batch_learned_holdout(fbh_batch* b, long count);
/* the network's mean loss on pairs [k0, k0 + count): cross-entropy against the conditional (kind 1) or the recorded
 * rule (kind 0); for kind 1 also the pairs' mean entropy and log allowed class count (either pointer may be null) */
int fbh_batch_learned_eval(fbh_batch* b, long k0, long count, double* mean_loss, double* mean_entropy,
                           double* mean_log_allowed);
/* copy pairs [k0, k0 + count) to the host (any pointer may be null): features [count][513], ranges [count][32]
 * (L | U << 8), cls [count] (v * 255 + c), entropy [count][2] (kind 1: the conditional's entropy and the log of the
 * allowed class count; kind 0: NaN) */
int fbh_batch_learned_pairs_copy(fbh_batch* b, long k0, long count, float* features, uint16_t* ranges, int* cls,
                                 float* entropy);
/* steps Adam steps on minibatches of the collected pairs (device-resident; soft targets for kind 1), loss_out [steps]
 * may be null */
int fbh_batch_learned_train(fbh_batch* b, int steps, int batch, float lr, unsigned seed, float* loss_out);
/* average duration (ms) of the learned sweeps so far, and their count */
int fbh_batch_learned_timing(const fbh_batch* b, double* avg_ms, int* nsweeps);

/* ---------------------------------------------------------------- diagnostics */
/* Evaluate the device normal draw, logf/cosf tables and double math on host-given inputs (parity tests):
 *   kind 0: out_f[i] = normal from raw generator outputs (in_u[2i], in_u[2i+1])
 *   kind 1: out_f[i] = sqrtf(-2*logf(u)) for u index in_u[i] (top 24 bits)
 *   kind 2: out_d[i] = log(in_d[i]); kind 3: out_d[i] = exp(in_d[i]); kind 4: out_d[i] = sqrt(in_d[i])
 *   kind 5: out_d[i] = in_d[2i] / in_d[2i+1] */
int fbh_debug_math(int kind, const uint32_t* in_u, const double* in_d, long count, float* out_f, double* out_d,
                   int device);
/* host libm image entries [k0, k0+count) of table kind 0 (logf), 1 (cosf of 6.28318531f*u), 2 (log of the double u)
 * on the generator grid u_k = (k + 1/2) 2^-24; out is float for kinds 0-1, double for kind 2 (no device needed) */
int fbh_debug_libm_host(int kind, long k0, long count, void* out);
/* per-tree-update trace of the sampler: cap > 0 allocates cap records per chain for the following launches;
 * cap = 0 copies chain's records (struct layout: fbh::TraceRec) into host_out */
int fbh_batch_trace(fbh_batch* b, int cap, void* host_out, int chain);
/* one launch of a plain copy kernel over two fresh device buffers of mib MiB (1..16384); its duration in ms goes to
 * *ms. Used to calibrate hardware byte counters (FETCH_SIZE / WRITE_SIZE) against a transfer of known size. */
int fbh_debug_copy(long mib, int device, double* ms);
/* Host only: the generator state 128 blocks steps (64 blocks attempt slots) after state[4], through the jump tables
 * of the device-wide probit redraw (the powers T^(128 2^k) of the state transition); blocks < 2^20. */
int fbh_debug_jump(const uint32_t* state, unsigned long long blocks, uint32_t* out);
/* the exact split conditional of fbh_batch_learned_target kind 1 at one node given on the host (the device function
 * the pairs use): n rows with cutpoint ranks q [p][n], membership member[n] (1 = the row is in the node), weights w and
 * partial residuals r; ranges [p] = L | U << 8 of the node's allowed cuts (U < L: none); the node's depth and
 * good-variable count, the tree's good-leaf count ngood, the depth prior table pg [64], sigma and tau. q_out [p][255]
 * = q(v, c); *entropy = its entropy (-1 when no rule qualifies); cls_out[t] = the class v * 255 + c at the inverse CDF
 * of u[t] (classes in that order). No reference counterpart. */
int fbh_debug_split_conditional(int n, int p, const uint8_t* q, const uint8_t* member, const double* w,
                                const double* r, const uint16_t* ranges, int depth, int ngv, int ngood,
                                const double* pg, double sigma, double tau, const double* u, int nu, float* q_out,
                                int* cls_out, double* entropy);
/* library build information */
const char* fbh_version(void);

#ifdef __cplusplus
}
#endif
#endif
