// The learned birth proposal's network on gfx950: a 3-layer MLP (d_in -> h1 -> h2 -> d_out logits, ReLU), BF16
// weights and activations, FP32 accumulation on the matrix cores (v_mfma_f32_32x32x16_bf16), FP32 master weights and
// Adam. Forward, masked-softmax cross-entropy, backward and the update are all device kernels; the host only orders
// them. (No counterpart in the reference: the proposal it replaces is BART's uniform rule draw, src/bart_kernel.cpp:433-437.)
//
// Layout: weights W_l are [in][out] row-major (bf16 copy + fp32 master), batches are [rows][features] row-major. Every
// dimension is padded: rows to 64, features to 64; padding is zero and stays zero (its gradients are zero).
#include <cmath>
#include <cstring>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kBM = 64, kBN = 64, kBK = 64;  // block tile; 4 waves of 32 x 32, four MFMAs per wave and k-step
constexpr int kLdsK = kBK + 8;                // k stride in LDS (bf16 elements), keeps 16-byte alignment

enum Epi { EPI_RELU_BF16 = 0, EPI_BIAS_F32 = 1, EPI_F32 = 2, EPI_DRELU = 3 };

// C[M x N] = A[M x K] * B[K x N] (+ epilogue). A(m,k) = TA ? A[k*lda + m] : A[m*lda + k];
// B(k,n) = TB ? B[n*ldb + k] : B[k*ldb + n]. M, N multiples of 64, K a multiple of 64.
//   EPI_RELU_BF16: Cb = bf16(relu(acc + bias))        EPI_BIAS_F32: Cf = acc + bias
//   EPI_F32:       Cf = acc                           EPI_DRELU: g = acc * (aux > 0): Cf = g, Cb = bf16(g)
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const bf16* __restrict__ A, const bf16* __restrict__ B, int K, int lda,
                                              int ldb, float* __restrict__ Cf, bf16* __restrict__ Cb, int ldc,
                                              const float* __restrict__ bias, const bf16* __restrict__ aux) {
  __shared__ __align__(16) bf16 As[kBM][kLdsK];
  __shared__ __align__(16) bf16 Bs[kBN][kLdsK];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc = {};
  for (int k0 = 0; k0 < K; k0 += kBK) {
    // global -> LDS: 64 x 64 of each operand, 16 bf16 per thread (two 8-wide pieces), stored [row][k]
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int t = tid + 256 * h;
      if (!TA) {
        const int r = t >> 3, kc = (t & 7) * 8;
        *reinterpret_cast<uint4*>(&As[r][kc]) =
            *reinterpret_cast<const uint4*>(A + (size_t)(m0 + r) * lda + k0 + kc);
      } else {
        const int kr = t >> 3, mc = (t & 7) * 8;
        const uint4 v = *reinterpret_cast<const uint4*>(A + (size_t)(k0 + kr) * lda + m0 + mc);
        const bf16* e = reinterpret_cast<const bf16*>(&v);
#pragma unroll
        for (int u = 0; u < 8; u++) As[mc + u][kr] = e[u];
      }
      if (TB) {
        const int r = t >> 3, kc = (t & 7) * 8;
        *reinterpret_cast<uint4*>(&Bs[r][kc]) =
            *reinterpret_cast<const uint4*>(B + (size_t)(n0 + r) * ldb + k0 + kc);
      } else {
        const int kr = t >> 3, nc = (t & 7) * 8;
        const uint4 v = *reinterpret_cast<const uint4*>(B + (size_t)(k0 + kr) * ldb + n0 + nc);
        const bf16* e = reinterpret_cast<const bf16*>(&v);
#pragma unroll
        for (int u = 0; u < 8; u++) Bs[nc + u][kr] = e[u];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 16) {
      // lane l: row / column l % 32, k = kk + 8 (l / 32) .. + 8
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + (lane & 31)][kk + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + (lane & 31)][kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  // acc[4 i + j] holds C[8 i + 4 (l / 32) + j][l % 32] of the wave's 32 x 32 tile
  const int n = n0 + wn + (lane & 31);
  const float bn = (EPI == EPI_RELU_BF16 || EPI == EPI_BIAS_F32) ? bias[n] : 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int m = m0 + wm + 8 * i + 4 * (lane >> 5) + j;
      const float v = acc[4 * i + j];
      const size_t o = (size_t)m * ldc + n;
      if (EPI == EPI_RELU_BF16) {
        Cb[o] = (bf16)fmaxf(v + bn, 0.f);
      } else if (EPI == EPI_BIAS_F32) {
        Cf[o] = v + bn;
      } else if (EPI == EPI_F32) {
        Cf[o] = v;
      } else {
        const float g = (float)aux[o] > 0.f ? v : 0.f;
        Cf[o] = g;
        Cb[o] = (bf16)g;
      }
    }
  }
}

// Masked softmax cross-entropy of one row per block: loss[m] = -log softmax(z_valid)[target]; dZ = (softmax - onehot) *
// scale over the valid entries, 0 elsewhere (rows >= M are all zero). SOFT: the target is a distribution q [M][d_out]
// over the valid entries instead of a class: loss[m] = sum_n q_n (lse - z_n), dZ = (sum(q) softmax - q) * scale.
template <bool SOFT>
__global__ __launch_bounds__(256) void k_softmax_xent(const float* __restrict__ Z, int ldz, int d_out, int M,
                                                      const int* __restrict__ target, const uint8_t* __restrict__ mask,
                                                      const float* __restrict__ soft, float scale,
                                                      float* __restrict__ dZ, bf16* __restrict__ dZb,
                                                      float* __restrict__ loss) {
  __shared__ float red[256];
  const int m = blockIdx.x, tid = threadIdx.x;
  const float* z = Z + (size_t)m * ldz;
  float* dz = dZ + (size_t)m * ldz;
  bf16* dzb = dZb + (size_t)m * ldz;
  if (m >= M) {
    for (int n = tid; n < ldz; n += 256) {
      dz[n] = 0.f;
      dzb[n] = (bf16)0.f;
    }
    return;
  }
  const uint8_t* mk = mask ? mask + (size_t)m * d_out : nullptr;
  float mx = -INFINITY;
  for (int n = tid; n < d_out; n += 256)
    if (!mk || mk[n]) mx = fmaxf(mx, z[n]);
  red[tid] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] = fmaxf(red[tid], red[tid + s]);
    __syncthreads();
  }
  mx = red[0];
  __syncthreads();
  float se = 0.f;
  for (int n = tid; n < d_out; n += 256)
    if (!mk || mk[n]) se += expf(z[n] - mx);
  red[tid] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  se = red[0];
  const float lse = mx + logf(se);
  if constexpr (SOFT) {
    const float* q = soft + (size_t)m * d_out;
    __syncthreads();
    float sq = 0.f, sl = 0.f;
    for (int n = tid; n < d_out; n += 256)
      if ((!mk || mk[n]) && q[n] > 0.f) {
        sq += q[n];
        sl += q[n] * (lse - z[n]);
      }
    red[tid] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    sq = red[0];
    __syncthreads();
    red[tid] = sl;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    sl = red[0];
    for (int n = tid; n < ldz; n += 256) {
      float g = 0.f;
      if (n < d_out && (!mk || mk[n])) g = (sq * expf(z[n] - lse) - q[n]) * scale;
      dz[n] = g;
      dzb[n] = (bf16)g;
    }
    if (tid == 0) loss[m] = sl;
  } else {
    const int t = target[m];
    for (int n = tid; n < ldz; n += 256) {
      float g = 0.f;
      if (n < d_out && (!mk || mk[n])) g = (expf(z[n] - lse) - (n == t ? 1.f : 0.f)) * scale;
      dz[n] = g;
      dzb[n] = (bf16)g;
    }
    if (tid == 0) loss[m] = lse - z[t];
  }
}

// out[n] = sum over rows m < M of X[m][n] (bias gradients): a block takes 32 columns, its 8 row slices (threads
// 32 s .. 32 s + 31) sum rows s, s + 8, ... of them, and the slices are added in a fixed order
constexpr int kColW = 32, kColS = 8;
__global__ __launch_bounds__(kColW * kColS) void k_colsum(const float* __restrict__ X, int M, int ld, int N,
                                                          float* __restrict__ out) {
  __shared__ float part[kColS][kColW];
  const int c = threadIdx.x % kColW, sl = threadIdx.x / kColW;
  const int n = blockIdx.x * kColW + c;
  float s = 0.f;
  if (n < N)
    for (int m = sl; m < M; m += kColS) s += X[(size_t)m * ld + n];
  part[sl][c] = s;
  __syncthreads();
  if (sl == 0 && n < N) {
    float t = part[0][c];
    for (int k = 1; k < kColS; k++) t += part[k][c];
    out[n] = t;
  }
}

// Adam on fp32 masters (Kingma & Ba 2015), bf16 copies refreshed (wb [rows][cols], and its transpose wbt when given);
// bias-corrected step size passed in
__global__ void k_adam(float* __restrict__ w, float* __restrict__ mo, float* __restrict__ ve,
                       const float* __restrict__ g, bf16* __restrict__ wb, long count, float step, float b1, float b2,
                       float eps, bf16* __restrict__ wbt = nullptr, int rows = 0, int cols = 0) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float m = b1 * mo[i] + (1.f - b1) * gi;
  const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
  mo[i] = m;
  ve[i] = v;
  const float nw = w[i] - step * m / (sqrtf(v) + eps);
  w[i] = nw;
  if (wb) wb[i] = (bf16)nw;
  if (wbt) wbt[(i % cols) * (long)rows + i / cols] = (bf16)nw;
}

// x [rows][cols] fp32 -> y (same layout) and yt [cols][rows] in bf16
__global__ void k_to_bf16(const float* __restrict__ x, bf16* __restrict__ y, bf16* __restrict__ yt, int rows,
                          int cols) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)rows * cols) return;
  const bf16 v = (bf16)x[i];
  y[i] = v;
  yt[(i % cols) * (long)rows + i / cols] = v;
}

// rows of x [M][d_in] (fp32, device) into the zero-padded bf16 batch [Mp][ld]
__global__ void k_pack_rows(const float* __restrict__ x, int M, int d_in, bf16* __restrict__ xb, int Mp, int ld) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)Mp * ld) return;
  const int m = (int)(i / ld), k = (int)(i % ld);
  xb[i] = (bf16)((m < M && k < d_in) ? x[(size_t)m * d_in + k] : 0.f);
}

}  // namespace fbh

using namespace fbh;

namespace {
inline int up64(int x) { return (x + 63) / 64 * 64; }

// deterministic He-normal initialisation (splitmix64 + Box-Muller in double)
struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double unif() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double normal() { return std::sqrt(-2.0 * std::log(unif())) * std::cos(6.283185307179586 * unif()); }
};
}  // namespace

struct fbh_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  int d[4] = {0, 0, 0, 0};   // d_in, h1, h2, d_out
  int p[4] = {0, 0, 0, 0};   // padded
  int maxM = 0, Mp = 0;
  long wcount[3] = {0, 0, 0}, bcount[3] = {0, 0, 0};
  // parameters: fp32 masters, Adam moments, bf16 copies, gradients (padded layouts)
  float *W[3] = {}, *Bv[3] = {}, *mW[3] = {}, *vW[3] = {}, *mB[3] = {}, *vB[3] = {}, *gW[3] = {}, *gB[3] = {};
  bf16* Wb[3] = {};
  bf16* WbT[3] = {};  // transposed bf16 copies [out][in]: the forward's weight operand, contiguous along k
  // activations
  bf16 *Xb = nullptr, *H1 = nullptr, *H2 = nullptr, *dZb = nullptr, *dH2b = nullptr, *dH1b = nullptr;
  float *Z = nullptr, *dZ = nullptr, *dH2 = nullptr, *dH1 = nullptr, *loss = nullptr, *xin = nullptr;
  int* tgt = nullptr;
  uint8_t* mask = nullptr;
  float* soft = nullptr;  // soft targets [maxM][d_out] (allocated on first use)
  int adam_t = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double fwd_ms = 0;
  int fwd_n = 0;
};

static void mlp_free(fbh_mlp* n) {
  if (!n) return;
  (void)hipSetDevice(n->device);
  for (int l = 0; l < 3; l++) {
    void* ps[] = {n->W[l], n->Bv[l], n->mW[l], n->vW[l], n->mB[l], n->vB[l], n->gW[l], n->gB[l], n->Wb[l], n->WbT[l]};
    for (void* q : ps)
      if (q) (void)hipFree(q);
  }
  void* ps[] = {n->Xb, n->H1, n->H2, n->dZb, n->dH2b, n->dH1b, n->Z, n->dZ, n->dH2, n->dH1, n->loss, n->xin,
                n->tgt, n->mask, n->soft};
  for (void* q : ps)
    if (q) (void)hipFree(q);
  if (n->e0) (void)hipEventDestroy(n->e0);
  if (n->e1) (void)hipEventDestroy(n->e1);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
}

template <class T>
static int zalloc(T** ptr, size_t count) {
  *ptr = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(ptr), std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    set_error(FBH_ENOMEM, "fbh_mlp: hipMalloc failed");
    return FBH_ENOMEM;
  }
  if (hipMemset(*ptr, 0, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    set_error(FBH_EHIP, "fbh_mlp: hipMemset failed");
    return FBH_EHIP;
  }
  return FBH_OK;
}

extern "C" long fbh_mlp_nparams(const fbh_mlp* n) {
  if (!n) return -1;
  return (long)n->d[0] * n->d[1] + n->d[1] + (long)n->d[1] * n->d[2] + n->d[2] + (long)n->d[2] * n->d[3] + n->d[3];
}

extern "C" int fbh_mlp_create(const int* dims, int max_batch, unsigned seed, int device, fbh_mlp** out) {
  if (!out || !dims || max_batch < 1 || max_batch > (1 << 20)) {
    set_error(FBH_EINVAL, "fbh_mlp_create: bad argument");
    return FBH_EINVAL;
  }
  *out = nullptr;
  for (int i = 0; i < 4; i++)
    if (dims[i] < 1 || dims[i] > 65536) {
      set_error(FBH_EINVAL, "fbh_mlp_create: dimensions must be in 1..65536");
      return FBH_EINVAL;
    }
  int rc = ensure_device(device);
  if (rc) return rc;
  fbh_mlp* n = new fbh_mlp();
  n->device = device;
  for (int i = 0; i < 4; i++) {
    n->d[i] = dims[i];
    n->p[i] = up64(dims[i]);
  }
  n->maxM = max_batch;
  n->Mp = up64(max_batch);
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&n->e0) ||
      hipEventCreate(&n->e1)) {
    set_error(FBH_EHIP, "fbh_mlp_create: stream / event creation failed");
    mlp_free(n);
    return FBH_EHIP;
  }
  for (int l = 0; l < 3; l++) {
    n->wcount[l] = (long)n->p[l] * n->p[l + 1];
    n->bcount[l] = n->p[l + 1];
    if ((rc = zalloc(&n->W[l], n->wcount[l])) || (rc = zalloc(&n->mW[l], n->wcount[l])) ||
        (rc = zalloc(&n->vW[l], n->wcount[l])) || (rc = zalloc(&n->gW[l], n->wcount[l])) ||
        (rc = zalloc(&n->Wb[l], n->wcount[l])) || (rc = zalloc(&n->WbT[l], n->wcount[l])) ||
        (rc = zalloc(&n->Bv[l], n->bcount[l])) ||
        (rc = zalloc(&n->mB[l], n->bcount[l])) || (rc = zalloc(&n->vB[l], n->bcount[l])) ||
        (rc = zalloc(&n->gB[l], n->bcount[l]))) {
      mlp_free(n);
      return rc;
    }
  }
  const size_t Mp = n->Mp;
  if ((rc = zalloc(&n->Xb, Mp * n->p[0])) || (rc = zalloc(&n->H1, Mp * n->p[1])) ||
      (rc = zalloc(&n->H2, Mp * n->p[2])) || (rc = zalloc(&n->Z, Mp * n->p[3])) ||
      (rc = zalloc(&n->dZ, Mp * n->p[3])) || (rc = zalloc(&n->dZb, Mp * n->p[3])) ||
      (rc = zalloc(&n->dH2, Mp * n->p[2])) || (rc = zalloc(&n->dH2b, Mp * n->p[2])) ||
      (rc = zalloc(&n->dH1, Mp * n->p[1])) || (rc = zalloc(&n->dH1b, Mp * n->p[1])) ||
      (rc = zalloc(&n->loss, Mp)) || (rc = zalloc(&n->xin, (size_t)max_batch * n->d[0])) ||
      (rc = zalloc(&n->tgt, Mp)) || (rc = zalloc(&n->mask, (size_t)max_batch * n->d[3]))) {
    mlp_free(n);
    return rc;
  }
  // He-normal weights, zero biases
  SplitMix sm{seed * 0x9E3779B97F4A7C15ull + 1};
  std::vector<float> params(fbh_mlp_nparams(n));
  size_t o = 0;
  for (int l = 0; l < 3; l++) {
    const double sd = std::sqrt(2.0 / n->d[l]);
    for (long i = 0; i < (long)n->d[l] * n->d[l + 1]; i++) params[o++] = (float)(sd * sm.normal());
    for (int i = 0; i < n->d[l + 1]; i++) params[o++] = 0.f;
  }
  *out = n;
  return fbh_mlp_params(n, nullptr, params.data());
}

// get (if get) and/or set (if set) the parameters, packed W1 [d_in][h1], b1, W2, b2, W3, b3 (fp32)
extern "C" int fbh_mlp_params(fbh_mlp* n, float* get, const float* set) {
  if (!n) {
    set_error(FBH_EINVAL, "fbh_mlp_params: null handle");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(n->device));
  FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
  size_t o = 0;
  for (int l = 0; l < 3; l++) {
    const int rows = n->d[l], cols = n->d[l + 1], ld = n->p[l + 1];
    std::vector<float> buf((size_t)n->wcount[l]);
    if (get) {
      FBH_HIP_CHECK(hipMemcpy(buf.data(), n->W[l], sizeof(float) * buf.size(), hipMemcpyDeviceToHost));
      for (int r = 0; r < rows; r++) std::memcpy(get + o + (size_t)r * cols, buf.data() + (size_t)r * ld, 4 * cols);
    }
    if (set) {
      std::fill(buf.begin(), buf.end(), 0.f);
      for (int r = 0; r < rows; r++) std::memcpy(buf.data() + (size_t)r * ld, set + o + (size_t)r * cols, 4 * cols);
      FBH_HIP_CHECK(hipMemcpy(n->W[l], buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n->wcount[l] + 255) / 256)), dim3(256), 0, n->stream, n->W[l],
                         n->Wb[l], n->WbT[l], n->p[l], n->p[l + 1]);
      FBH_HIP_CHECK(hipGetLastError());
    }
    o += (size_t)rows * cols;
    std::vector<float> bb((size_t)n->bcount[l], 0.f);
    if (get) {
      FBH_HIP_CHECK(hipMemcpy(bb.data(), n->Bv[l], sizeof(float) * bb.size(), hipMemcpyDeviceToHost));
      std::memcpy(get + o, bb.data(), 4 * cols);
    }
    if (set) {
      std::fill(bb.begin(), bb.end(), 0.f);
      std::memcpy(bb.data(), set + o, 4 * cols);
      FBH_HIP_CHECK(hipMemcpy(n->Bv[l], bb.data(), sizeof(float) * bb.size(), hipMemcpyHostToDevice));
    }
    o += cols;
  }
  if (set) {
    n->adam_t = 0;
    for (int l = 0; l < 3; l++) {
      FBH_HIP_CHECK(hipMemset(n->mW[l], 0, sizeof(float) * n->wcount[l]));
      FBH_HIP_CHECK(hipMemset(n->vW[l], 0, sizeof(float) * n->wcount[l]));
      FBH_HIP_CHECK(hipMemset(n->mB[l], 0, sizeof(float) * n->bcount[l]));
      FBH_HIP_CHECK(hipMemset(n->vB[l], 0, sizeof(float) * n->bcount[l]));
    }
  }
  FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
  return FBH_OK;
}

namespace fbh {
// forward of rows [0, M) of the device batch xin (fp32 [M][d_in]) into n->Z (fp32 logits [Mp][p3]): the input packed to
// the padded bf16 batch n->Xb (also the backward's first operand), then the three layers with the weights read through
// their transposed bf16 copies (contiguous along k, so every operand tile is loaded with 16-byte vectors)
int mlp_forward_dev(fbh_mlp* n, const float* xin_dev, int M, hipStream_t s) {
  const int Mp = up64(M);
  const int* p = n->p;
  hipLaunchKernelGGL(k_pack_rows, dim3((unsigned)(((long)Mp * p[0] + 255) / 256)), dim3(256), 0, s, xin_dev, M,
                     n->d[0], n->Xb, Mp, p[0]);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_RELU_BF16>), dim3(p[1] / kBN, Mp / kBM), dim3(256), 0, s, n->Xb,
                     n->WbT[0], p[0], p[0], p[0], nullptr, n->H1, p[1], n->Bv[0], nullptr);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_RELU_BF16>), dim3(p[2] / kBN, Mp / kBM), dim3(256), 0, s, n->H1,
                     n->WbT[1], p[1], p[1], p[1], nullptr, n->H2, p[2], n->Bv[1], nullptr);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_BIAS_F32>), dim3(p[3] / kBN, Mp / kBM), dim3(256), 0, s, n->H2,
                     n->WbT[2], p[2], p[2], p[2], n->Z, nullptr, p[3], n->Bv[2], nullptr);
  FBH_HIP_CHECK(hipGetLastError());
  return FBH_OK;
}

// the masked cross-entropy of rows [0, M) into n->loss and its gradient into n->dZ / n->dZb (after mlp_forward_dev);
// targets and masks in n->tgt / n->mask, or soft targets in n->soft
int mlp_loss_dev(fbh_mlp* n, int M, bool soft, hipStream_t s) {
  const int Mp = up64(M);
  const int* p = n->p;
  if (soft)
    hipLaunchKernelGGL(k_softmax_xent<true>, dim3(Mp), dim3(256), 0, s, n->Z, p[3], n->d[3], M, n->tgt, n->mask,
                       n->soft, 1.f / (float)M, n->dZ, n->dZb, n->loss);
  else
    hipLaunchKernelGGL(k_softmax_xent<false>, dim3(Mp), dim3(256), 0, s, n->Z, p[3], n->d[3], M, n->tgt, n->mask,
                       (const float*)nullptr, 1.f / (float)M, n->dZ, n->dZb, n->loss);
  FBH_HIP_CHECK(hipGetLastError());
  return FBH_OK;
}

// the soft-target buffer [max_batch][d_out] (allocated on first use)
float* mlp_soft_buffer(fbh_mlp* n) {
  if (!n->soft && hipMalloc(reinterpret_cast<void**>(&n->soft), sizeof(float) * (size_t)n->maxM * n->d[3]) !=
                      hipSuccess) {
    n->soft = nullptr;
    set_error(FBH_ENOMEM, "fbh_mlp: soft-target buffer allocation failed");
  }
  return n->soft;
}

// masked cross-entropy and the gradients of the mean loss over rows [0, M) (after mlp_forward_dev); targets and masks
// in n->tgt / n->mask (soft = false) or soft targets in n->soft
int mlp_backward_dev(fbh_mlp* n, int M, hipStream_t s, bool soft) {
  const int Mp = up64(M);
  const int* p = n->p;
  int rc = mlp_loss_dev(n, M, soft, s);
  if (rc) return rc;
  // layer 3
  hipLaunchKernelGGL((k_gemm<true, false, EPI_F32>), dim3(p[3] / kBN, p[2] / kBM), dim3(256), 0, s,
                     n->H2, n->dZb, Mp, p[2], p[3], n->gW[2], nullptr, p[3], nullptr, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3((p[3] + kColW - 1) / kColW), dim3(kColW * kColS), 0, s, n->dZ, Mp, p[3], p[3],
                     n->gB[2]);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_DRELU>), dim3(p[2] / kBN, Mp / kBM), dim3(256), 0, s,
                     n->dZb, n->Wb[2], p[3], p[3], p[3], n->dH2, n->dH2b, p[2], nullptr, n->H2);
  // layer 2
  hipLaunchKernelGGL((k_gemm<true, false, EPI_F32>), dim3(p[2] / kBN, p[1] / kBM), dim3(256), 0, s,
                     n->H1, n->dH2b, Mp, p[1], p[2], n->gW[1], nullptr, p[2], nullptr, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3((p[2] + kColW - 1) / kColW), dim3(kColW * kColS), 0, s, n->dH2, Mp, p[2], p[2],
                     n->gB[1]);
  hipLaunchKernelGGL((k_gemm<false, true, EPI_DRELU>), dim3(p[1] / kBN, Mp / kBM), dim3(256), 0, s,
                     n->dH2b, n->Wb[1], p[2], p[2], p[2], n->dH1, n->dH1b, p[1], nullptr, n->H1);
  // layer 1
  hipLaunchKernelGGL((k_gemm<true, false, EPI_F32>), dim3(p[1] / kBN, p[0] / kBM), dim3(256), 0, s,
                     n->Xb, n->dH1b, Mp, p[0], p[1], n->gW[0], nullptr, p[1], nullptr, nullptr);
  hipLaunchKernelGGL(k_colsum, dim3((p[1] + kColW - 1) / kColW), dim3(kColW * kColS), 0, s, n->dH1, Mp, p[1], p[1],
                     n->gB[0]);
  FBH_HIP_CHECK(hipGetLastError());
  return FBH_OK;
}

int mlp_adam_dev(fbh_mlp* n, float lr, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  n->adam_t++;
  const float step = lr * (float)(std::sqrt(1.0 - std::pow((double)b2, n->adam_t)) /
                                  (1.0 - std::pow((double)b1, n->adam_t)));
  for (int l = 0; l < 3; l++) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->wcount[l] + 255) / 256)), dim3(256), 0, s, n->W[l], n->mW[l],
                       n->vW[l], n->gW[l], n->Wb[l], n->wcount[l], step, b1, b2, eps, n->WbT[l], n->p[l],
                       n->p[l + 1]);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n->bcount[l] + 255) / 256)), dim3(256), 0, s, n->Bv[l], n->mB[l],
                       n->vB[l], n->gB[l], (bf16*)nullptr, n->bcount[l], step, b1, b2, eps, (bf16*)nullptr, 0, 0);
  }
  FBH_HIP_CHECK(hipGetLastError());
  return FBH_OK;
}

const float* mlp_logits(const fbh_mlp* n, int* ld) {
  *ld = n->p[3];
  return n->Z;
}
void mlp_dims(const fbh_mlp* n, int* d, int* max_batch) {
  for (int i = 0; i < 4; i++) d[i] = n->d[i];
  *max_batch = n->maxM;
}
// the network's own batch buffers (device): inputs [maxM][d_in], targets [maxM], masks [maxM][d_out], losses
void mlp_batch_buffers(fbh_mlp* n, float** x, int** tgt, uint8_t** mask, float** loss) {
  *x = n->xin;
  *tgt = n->tgt;
  *mask = n->mask;
  *loss = n->loss;
}
}  // namespace fbh

static int mlp_upload_batch(fbh_mlp* n, const float* x, const int* target, const uint8_t* mask, int M) {
  if (M < 1 || M > n->maxM) {
    set_error(FBH_EINVAL, "fbh_mlp: batch size outside 1..max_batch");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipMemcpyAsync(n->xin, x, sizeof(float) * (size_t)M * n->d[0], hipMemcpyHostToDevice, n->stream));
  if (target) {
    for (int m = 0; m < M; m++)
      if (target[m] < 0 || target[m] >= n->d[3] || (mask && !mask[(size_t)m * n->d[3] + target[m]])) {
        set_error(FBH_EINVAL, "fbh_mlp: target outside d_out or masked out");
        return FBH_EINVAL;
      }
    FBH_HIP_CHECK(hipMemcpyAsync(n->tgt, target, sizeof(int) * M, hipMemcpyHostToDevice, n->stream));
  }
  if (mask) {
    FBH_HIP_CHECK(hipMemcpyAsync(n->mask, mask, (size_t)M * n->d[3], hipMemcpyHostToDevice, n->stream));
  } else {
    FBH_HIP_CHECK(hipMemsetAsync(n->mask, 1, (size_t)M * n->d[3], n->stream));
  }
  return FBH_OK;
}

extern "C" int fbh_mlp_forward(fbh_mlp* n, const float* x, int M, float* logits) {
  if (!n || !x || !logits) {
    set_error(FBH_EINVAL, "fbh_mlp_forward: null argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(n->device));
  int rc = mlp_upload_batch(n, x, nullptr, nullptr, M);
  if (rc) return rc;
  FBH_HIP_CHECK(hipEventRecord(n->e0, n->stream));
  if ((rc = mlp_forward_dev(n, n->xin, M, n->stream))) return rc;
  FBH_HIP_CHECK(hipEventRecord(n->e1, n->stream));
  FBH_HIP_CHECK(hipMemcpy2DAsync(logits, sizeof(float) * n->d[3], n->Z, sizeof(float) * n->p[3],
                                 sizeof(float) * n->d[3], M, hipMemcpyDeviceToHost, n->stream));
  FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
  float ms = 0;
  if (hipEventElapsedTime(&ms, n->e0, n->e1) == hipSuccess) {
    n->fwd_ms += ms;
    n->fwd_n++;
  }
  return FBH_OK;
}

extern "C" int fbh_mlp_grad(fbh_mlp* n, const float* x, const int* target, const uint8_t* mask, int M, float* loss,
                            float* grads) {
  if (!n || !x || !target) {
    set_error(FBH_EINVAL, "fbh_mlp_grad: null argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(n->device));
  int rc = mlp_upload_batch(n, x, target, mask, M);
  if (rc) return rc;
  if ((rc = mlp_forward_dev(n, n->xin, M, n->stream)) || (rc = mlp_backward_dev(n, M, n->stream))) return rc;
  FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
  if (loss) {
    std::vector<float> l(M);
    FBH_HIP_CHECK(hipMemcpy(l.data(), n->loss, sizeof(float) * M, hipMemcpyDeviceToHost));
    double s = 0;
    for (float v : l) s += v;
    *loss = (float)(s / M);
  }
  if (grads) {
    size_t o = 0;
    for (int l = 0; l < 3; l++) {
      const int rows = n->d[l], cols = n->d[l + 1], ld = n->p[l + 1];
      std::vector<float> buf((size_t)n->wcount[l]);
      FBH_HIP_CHECK(hipMemcpy(buf.data(), n->gW[l], sizeof(float) * buf.size(), hipMemcpyDeviceToHost));
      for (int r = 0; r < rows; r++) std::memcpy(grads + o + (size_t)r * cols, buf.data() + (size_t)r * ld, 4 * cols);
      o += (size_t)rows * cols;
      FBH_HIP_CHECK(hipMemcpy(grads + o, n->gB[l], sizeof(float) * cols, hipMemcpyDeviceToHost));
      o += cols;
    }
  }
  return FBH_OK;
}

// `steps` Adam steps on minibatches of `batch` rows drawn (with a splitmix stream from `seed`) from the M-row dataset;
// loss_out (may be null) receives the mean loss of each step
extern "C" int fbh_mlp_train(fbh_mlp* n, const float* x, const int* target, const uint8_t* mask, int M, int steps,
                             int batch, float lr, unsigned seed, float* loss_out) {
  if (!n || !x || !target || M < 1 || steps < 0 || batch < 1 || batch > n->maxM || !(lr > 0.f)) {
    set_error(FBH_EINVAL, "fbh_mlp_train: bad argument");
    return FBH_EINVAL;
  }
  FBH_HIP_CHECK(hipSetDevice(n->device));
  SplitMix sm{(uint64_t)seed * 0xD1B54A32D192ED03ull + 7};
  std::vector<float> xb((size_t)batch * n->d[0]);
  std::vector<int> tb(batch);
  std::vector<uint8_t> mb(mask ? (size_t)batch * n->d[3] : 0);
  for (int st = 0; st < steps; st++) {
    for (int b = 0; b < batch; b++) {
      const int r = (int)(sm.next() % (uint64_t)M);
      std::memcpy(xb.data() + (size_t)b * n->d[0], x + (size_t)r * n->d[0], sizeof(float) * n->d[0]);
      tb[b] = target[r];
      if (mask) std::memcpy(mb.data() + (size_t)b * n->d[3], mask + (size_t)r * n->d[3], n->d[3]);
    }
    int rc = mlp_upload_batch(n, xb.data(), tb.data(), mask ? mb.data() : nullptr, batch);
    if (rc) return rc;
    if ((rc = mlp_forward_dev(n, n->xin, batch, n->stream)) || (rc = mlp_backward_dev(n, batch, n->stream)) ||
        (rc = mlp_adam_dev(n, lr, n->stream)))
      return rc;
    if (loss_out) {
      std::vector<float> l(batch);
      FBH_HIP_CHECK(hipMemcpyAsync(l.data(), n->loss, sizeof(float) * batch, hipMemcpyDeviceToHost, n->stream));
      FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
      double s = 0;
      for (float v : l) s += v;
      loss_out[st] = (float)(s / batch);
    }
  }
  FBH_HIP_CHECK(hipStreamSynchronize(n->stream));
  return FBH_OK;
}

extern "C" int fbh_mlp_timing(const fbh_mlp* n, double* fwd_ms, int* nfwd) {
  if (!n) return FBH_EINVAL;
  if (fwd_ms) *fwd_ms = n->fwd_n ? n->fwd_ms / n->fwd_n : 0;
  if (nfwd) *nfwd = n->fwd_n;
  return FBH_OK;
}

extern "C" void fbh_mlp_destroy(fbh_mlp* n) { mlp_free(n); }
