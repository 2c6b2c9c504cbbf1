// Internal declarations shared by the HIP translation units of libfastbart_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/fastbart_hip.h"

namespace fbh {

constexpr int kNodes = FBH_NODES;
constexpr int kLanes = 16;        // summation lanes of the reference (src/bart_kernel.h:25-27): row i -> lane i % 16
constexpr int kMaxDepth = 24;     // src/bart_kernel.h:33
constexpr int kMaxRank = kNodes / 2 + 2;  // at most 32 leaves + the two birth-split halves

// per-thread error text (fbh_last_error)
void set_error(int code, const std::string& msg);
int last_code();

#define FBH_HIP_CHECK(expr)                                                                   \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess) {                                                                   \
      ::fbh::set_error(FBH_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
      return FBH_EHIP;                                                                        \
    }                                                                                         \
  } while (0)

// One tree: 64 node slots (src/bart_kernel.cpp:112-116). Slot fields stay as the sampler left them, stale values
// included, because the output words carry them.
struct TreeDev {
  uint8_t l[kNodes], r[kNodes], par[kNodes], v[kNodes], c[kNodes], d[kNodes], ngv[kNodes];
  float th[kNodes];
  uint64_t used, leaf;
};
static_assert(sizeof(TreeDev) == 720, "TreeDev layout");

// Read-only description of one chain in the batch pools.
struct ChainDesc {
  int n, p;
  float tau, lambda, nu, sigest, offset;
  int root_ngv;
  int nc[FBH_PMAX];
  uint64_t q_off;     // bytes into the q pool ([p][n] u8)
  uint64_t row_off;   // elements into the row pools (Y, W, AF, TM, ybin)
  uint64_t leaf_off;  // bytes into the leaf-id pool ([m][n] u8)
  uint64_t tree_off;  // TreeDev index (chain * m)
  uint64_t sig_off;   // float index into sigma_out
  uint64_t tout_off;  // word index into trees_out ([ndpost][m][64])
  // row split: this workgroup's chain is member `member` of `split` workgroups sharing logical chain `group`, whose
  // rows number n_total (n is this member's share); keep = 1 if it writes the kept trees
  int n_total, split, member, group, keep;
  int ns;  // stride of the q and leaf-id columns: n rounded up to a multiple of 16 (16-byte aligned columns)
  int pad_[2];
};

// Mutable per-chain sampler scalars carried between launches.
struct ChainState {
  uint32_t rng[4];
  double sigma;
  int it, n_birth, n_death, n_cap;
  uint32_t xseq;  // row split: exchanges done so far
  int xerr;       // row split: an exchange timed out (the results are invalid)
};

// Device images of the host libm on the generator's 2^24-point uniform grid u_k = (k + 1/2) 2^-24:
//   logf_t[k] = logf(u_k), cosf_t[k] = cosf(6.28318531f * u_k), logd_t[k] = log((double)u_k).
// The reference's draws use the platform libm on exactly these arguments (src/bart_kernel.cpp:60, :289, :482, :80),
// and glibc's logf/cosf are not correctly rounded everywhere on this grid, so the device reads the host libm's own
// values. Built once per process, uploaded once per device.
struct LibmTables {
  const float* logf_t;
  const float* cosf_t;
  const double* logd_t;
};
int libm_tables(int device, LibmTables* out);

// Per-tree-update record for debugging / kernel-level parity (optional).
struct TraceRec {
  int32_t it, j, mode, nx, v, c, nl, nr, accepted, ngood, nnogs, cap;
  double pr, lalpha, lu, sl0, sl1, sr0, sr1;
  uint32_t rng_after[4];
};

int ensure_device(int device);

// the proposal network's device-level steps (fbh_mlp.hip), on a caller's stream
int mlp_forward_dev(fbh_mlp* n, const float* xin_dev, int M, hipStream_t s);
int mlp_backward_dev(fbh_mlp* n, int M, hipStream_t s, bool soft = false);
int mlp_loss_dev(fbh_mlp* n, int M, bool soft, hipStream_t s);
float* mlp_soft_buffer(fbh_mlp* n);
int mlp_adam_dev(fbh_mlp* n, float lr, hipStream_t s);
const float* mlp_logits(const fbh_mlp* n, int* ld);
void mlp_dims(const fbh_mlp* n, int* d, int* max_batch);
void mlp_batch_buffers(fbh_mlp* n, float** x, int** tgt, uint8_t** mask, float** loss);

}  // namespace fbh
