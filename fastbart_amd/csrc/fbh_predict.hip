// Output formats and prediction: treedraws text and varcount (src/glue.cpp:29-51), per-draw training fits
// (src/glue.cpp:52-56) on the device, and compile_trees / predict_draw (src/predict.cpp) with a tree-outer device
// kernel over flat child-pair arrays.
#include <atomic>
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

// ------------------------------------------------------------------ training fits from packed words
// One thread per (draw, row); the m trees of the draw are added in order to a double that starts at 0.0, each tree
// walked with q <= c going left. The draw's words are read through the cache (every thread of the draw reads the same
// words).
__global__ __launch_bounds__(256) void k_yhat(const uint64_t* __restrict__ trees, const uint8_t* __restrict__ q, int n,
                                              int ldq, int m, double* __restrict__ out) {
  const int dr = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t* T0 = trees + (size_t)dr * m * kNodes;
  double f = 0.0;
  for (int j = 0; j < m; j++) {
    const uint64_t* T = T0 + (size_t)j * kNodes;
    uint64_t w = T[0];
    for (int step = 0; step < kNodes && ((w >> 16) & 255); step++) {
      const int v = (int)(w & 255), c = (int)((w >> 8) & 255);
      const int s = q[(size_t)v * ldq + i] <= c ? (int)((w >> 16) & 255) : (int)((w >> 24) & 255);
      w = T[s];
    }
    f += (double)__uint_as_float((uint32_t)(w >> 32));
  }
  out[(size_t)dr * n + i] = f;
}

// ------------------------------------------------------------------ predictor kernels
// Rank form (the fast path). With the cutpoints xi[v][0..K-1] of a variable nondecreasing and free of NaN, the
// predicate x >= xi[v][c] of src/predict.cpp:68 is monotone in c, so it holds exactly when c < rank(x), where
// rank(x) = #{k : x >= xi[v][k]} (0 for NaN, which compares false with everything). A row is then described by p
// 16-bit ranks instead of p doubles, and a node by 8 bytes: the index of its left child (its own index for a leaf),
// the variable and the cut index (0xFFFF for a leaf, which no rank exceeds, so a leaf maps to itself). A tree is
// walked a fixed number of steps, its depth, with no branch: k = next[k] + (rank[var[k]] > c[k]).
constexpr int kRankThreads = 256;
constexpr int kRankTrees = 4;             // trees walked at once per row (independent walks, added in tree order)
constexpr int kRankRowsMax = 2 * kRankThreads;  // rows per workgroup at most (2 per thread)
constexpr uint32_t kLeafCut = 0xFFFFu;

// ranks [p][n] u16 of the rows x [p][n] against the cutpoints xi (flat, variable v < pcut at xi[off[v] ..
// off[v+1]-1]; a variable without cutpoints has rank 0): a binary search with the reference's own comparison x >= xi[k]
__global__ __launch_bounds__(256) void k_rank(const double* __restrict__ x, const double* __restrict__ xi,
                                              const int* __restrict__ off, int n, int p, int pcut,
                                              uint16_t* __restrict__ rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = blockIdx.y;
  if (i >= n || v >= p) return;
  const double xv = x[(size_t)v * n + i];
  const double* c = v < pcut ? xi + off[v] : xi;
  int lo = 0, hi = v < pcut ? off[v + 1] - off[v] : 0;  // invariant: x >= c[k] for k < lo, not for k >= hi
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (xv >= c[mid]) lo = mid + 1;
    else hi = mid;
  }
  rank[(size_t)v * n + i] = (uint16_t)lo;
}

// One workgroup = one draw x (256 R) rows. LDS holds the draw's node words and leaf values, its per-tree
// (root, depth) words and the rows' ranks [p][256 R]; every thread adds the trees in order for its R rows, from 0.0,
// then mu (src/predict.cpp:64-72). The trees are walked kRankTrees at a time (for the deepest of them; a leaf maps to
// itself), so that a thread has R x kRankTrees independent LDS chains in flight; their leaf values are then added in
// tree order.
template <int R>
__global__ __launch_bounds__(kRankThreads) void k_predict_rank(const uint64_t* __restrict__ node,
                                                               const double* __restrict__ theta,
                                                               const uint32_t* __restrict__ tinfo,
                                                               const int* __restrict__ draw_lo,
                                                               const int* __restrict__ draws,
                                                               const uint16_t* __restrict__ rank, int n, int p,
                                                               int ntree, int ndpost, double mu,
                                                               double* __restrict__ out) {
  constexpr int kRows = kRankThreads * R;
  extern __shared__ __align__(16) unsigned char pred_lds[];
  const int k = blockIdx.y;
  const int dr = draws[k] - 1;
  const int row0 = blockIdx.x * kRows;
  if (dr < 0 || dr >= ndpost) {  // a draw outside 1..ndpost (device draw lists are not checked on the host): NaN
    for (int r = threadIdx.x; r < kRows; r += kRankThreads)
      if (row0 + r < n) out[(size_t)k * n + row0 + r] = __builtin_nan("");
    return;
  }
  const int lo = draw_lo[dr], cnt = draw_lo[dr + 1] - lo;
  uint64_t* s_node = reinterpret_cast<uint64_t*>(pred_lds);
  double* s_theta = reinterpret_cast<double*>(s_node + cnt);
  uint32_t* s_tinfo = reinterpret_cast<uint32_t*>(s_theta + cnt);
  uint16_t* s_rank = reinterpret_cast<uint16_t*>(s_tinfo + ((ntree + 3) & ~3));
  for (int t = threadIdx.x; t < cnt; t += kRankThreads) {
    s_node[t] = node[lo + t];
    s_theta[t] = theta[lo + t];
  }
  for (int t = threadIdx.x; t < ntree; t += kRankThreads) s_tinfo[t] = tinfo[(size_t)dr * ntree + t];
  for (int v = 0; v < p; v++)
    for (int r = threadIdx.x; r < kRows; r += kRankThreads) {
      const int i = row0 + r;
      s_rank[v * kRows + r] = i < n ? rank[(size_t)v * n + i] : (uint16_t)0;
    }
  __syncthreads();
  double f[R];
  const uint16_t* rk[R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    f[j] = 0.0;
    rk[j] = s_rank + threadIdx.x + j * kRankThreads;
  }
  for (int t0 = 0; t0 < ntree; t0 += kRankTrees) {
    // roots and the deepest depth of the group; trees past the end repeat tree t0 and are not added
    int root[kRankTrees];
    int depth = 0;
#pragma unroll
    for (int u = 0; u < kRankTrees; u++) {
      const uint32_t ti = s_tinfo[t0 + u < ntree ? t0 + u : t0];
      root[u] = (int)(ti & 0xFFFFFFu);
      depth = max(depth, (int)(ti >> 24));
    }
    int kk[R][kRankTrees];
#pragma unroll
    for (int j = 0; j < R; j++)
#pragma unroll
      for (int u = 0; u < kRankTrees; u++) kk[j][u] = root[u];
    for (int d = 0; d < depth; d++) {
#pragma unroll
      for (int j = 0; j < R; j++)
#pragma unroll
        for (int u = 0; u < kRankTrees; u++) {
          const uint64_t w = s_node[kk[j][u]];
          const uint32_t var = (uint32_t)(w >> 32) & 0xFFFFu, c = (uint32_t)(w >> 48);
          kk[j][u] = (int)(uint32_t)w + (int)((uint32_t)rk[j][var * kRows] > c);
        }
    }
#pragma unroll
    for (int u = 0; u < kRankTrees; u++)
      if (t0 + u < ntree) {
#pragma unroll
        for (int j = 0; j < R; j++) f[j] += s_theta[kk[j][u]];
      }
  }
#pragma unroll
  for (int j = 0; j < R; j++) {
    const int i = row0 + threadIdx.x + j * kRankThreads;
    if (i < n) out[(size_t)k * n + i] = f[j] + mu;
  }
}

// General form: any cutpoint table, any draw size. Tree-outer over a block of rows: the workgroup stages the flat node
// arrays of one draw in LDS (when they fit) and every thread adds the draw's trees in order for its row; x >= cut goes
// to the right child (src/predict.cpp:64-72).
constexpr int kPredThreads = 256;
constexpr int kPredLdsNodes = 2048;  // 2048 nodes x (int var, int child, double cut, double theta) = 48 KiB

__global__ __launch_bounds__(kPredThreads) void k_predict(const int* __restrict__ var, const int* __restrict__ child,
                                                          const double* __restrict__ cut,
                                                          const double* __restrict__ theta,
                                                          const int* __restrict__ root, const int* __restrict__ draws,
                                                          const int* __restrict__ draw_lo, const double* __restrict__ x,
                                                          int n, int ntree, int ndpost, double mu,
                                                          double* __restrict__ out) {
  __shared__ int s_var[kPredLdsNodes];
  __shared__ int s_child[kPredLdsNodes];
  __shared__ double s_cut[kPredLdsNodes];
  __shared__ double s_theta[kPredLdsNodes];
  const int k = blockIdx.y;
  const int dr = draws[k] - 1;
  if (dr < 0 || dr >= ndpost) {  // a draw outside 1..ndpost (device draw lists are not checked on the host): NaN
    const int i = blockIdx.x * kPredThreads + threadIdx.x;
    if (i < n) out[(size_t)k * n + i] = __builtin_nan("");
    return;
  }
  const int* r0 = root + (size_t)dr * ntree;
  const int lo = draw_lo[dr], hi = draw_lo[dr + 1];
  const int cnt = hi - lo;
  const bool staged = cnt <= kPredLdsNodes;
  if (staged) {
    for (int t = threadIdx.x; t < cnt; t += kPredThreads) {
      s_var[t] = var[lo + t];
      s_child[t] = child[lo + t] >= 0 ? child[lo + t] - lo : -1;
      s_cut[t] = cut[lo + t];
      s_theta[t] = theta[lo + t];
    }
  }
  __syncthreads();
  const int i = blockIdx.x * kPredThreads + threadIdx.x;
  if (i >= n) return;
  double f = 0.0;
  if (staged) {
    for (int t = 0; t < ntree; t++) {
      int kk = r0[t] - lo;
      while (s_child[kk] >= 0) kk = s_child[kk] + (int)(x[(size_t)s_var[kk] * n + i] >= s_cut[kk]);
      f += s_theta[kk];
    }
  } else {
    for (int t = 0; t < ntree; t++) {
      int kk = r0[t];
      while (child[kk] >= 0) kk = child[kk] + (int)(x[(size_t)var[kk] * n + i] >= cut[kk]);
      f += theta[kk];
    }
  }
  out[(size_t)k * n + i] = f + mu;
}

}  // namespace fbh

using namespace fbh;

// ==================================================================== yhat
namespace fbh {
// q [p][ldq] (column v at q + v ldq)
int yhat_device(const uint64_t* d_trees, const uint8_t* d_q, int n, int ldq, int ndpost, int m, hipStream_t stream,
                double* host_out) {
  if (ndpost == 0 || n == 0) return FBH_OK;
  double* d_out = nullptr;
  FBH_HIP_CHECK(hipMalloc(&d_out, sizeof(double) * (size_t)ndpost * n));
  dim3 grid((n + 255) / 256, ndpost);
  hipLaunchKernelGGL(k_yhat, grid, dim3(256), 0, stream, d_trees, d_q, n, ldq, m, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess)
    e = hipMemcpy(host_out, d_out, sizeof(double) * (size_t)ndpost * n, hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  if (e != hipSuccess) {
    set_error(FBH_EHIP, std::string("fbh_yhat: ") + hipGetErrorString(e));
    return FBH_EHIP;
  }
  return FBH_OK;
}
}  // namespace fbh

extern "C" int fbh_yhat(const uint64_t* trees, const uint8_t* q, int n, int p, int ndpost, int m, int device,
                        double* yhat) {
  if (!trees || !q || !yhat || n < 1 || p < 1 || ndpost < 0 || m < 1) {
    set_error(FBH_EINVAL, "fbh_yhat: bad argument");
    return FBH_EINVAL;
  }
  int rc = ensure_device(device);
  if (rc) return rc;
  // the walk indexes q only with variables taken from internal nodes; check them on the host first
  const size_t nw = (size_t)ndpost * m * kNodes;
  for (size_t w = 0; w < nw; w++) {
    const uint64_t x = trees[w];
    if (((x >> 16) & 255) && (int)(x & 255) >= p) {
      set_error(FBH_EINVAL, "fbh_yhat: an internal node splits on a variable >= p");
      return FBH_EINVAL;
    }
  }
  uint64_t* dt = nullptr;
  uint8_t* dq = nullptr;
  FBH_HIP_CHECK(hipMalloc(&dt, sizeof(uint64_t) * (nw ? nw : 1)));
  FBH_HIP_CHECK(hipMalloc(&dq, (size_t)p * n));
  FBH_HIP_CHECK(hipMemcpy(dt, trees, sizeof(uint64_t) * nw, hipMemcpyHostToDevice));
  FBH_HIP_CHECK(hipMemcpy(dq, q, (size_t)p * n, hipMemcpyHostToDevice));
  rc = yhat_device(dt, dq, n, n, ndpost, m, nullptr, yhat);
  (void)hipFree(dt);
  (void)hipFree(dq);
  return rc;
}

// ==================================================================== treedraws text
namespace {
// one tree in pre-order (ids root 1, children 2id / 2id+1), src/glue.cpp:40-51; %.10g for the leaf values
void tree_lines(const uint64_t* T, int p, std::string& s, int* vc) {
  int st_s[kNodes + 2];
  uint64_t st_id[kNodes + 2];
  int ord_s[kNodes];
  uint64_t ord_id[kNodes];
  int top = 0, cnt = 0;
  st_s[top] = 0;
  st_id[top++] = 1;
  while (top > 0 && cnt < kNodes) {
    --top;
    const int sl = st_s[top];
    const uint64_t id = st_id[top];
    ord_s[cnt] = sl;
    ord_id[cnt++] = id;
    const int L = (int)((T[sl] >> 16) & 255), R = (int)((T[sl] >> 24) & 255);
    if (L) {
      st_s[top] = R;
      st_id[top++] = 2 * id + 1;
      st_s[top] = L;
      st_id[top++] = 2 * id;
    }
  }
  char buf[96];
  auto put_int = [&](unsigned long long v) {
    auto r = std::to_chars(buf, buf + sizeof buf, v);
    s.append(buf, r.ptr);
  };
  put_int((unsigned long long)cnt);
  s.push_back('\n');
  for (int k = 0; k < cnt; k++) {
    const int sl = ord_s[k];
    const bool leaf = ((T[sl] >> 16) & 255) == 0;
    const int v = leaf ? 0 : (int)(T[sl] & 255), c = leaf ? 0 : (int)((T[sl] >> 8) & 255);
    if (!leaf && vc && v < p) vc[v] += 1;
    double th = 0.0;
    if (leaf) {
      const uint32_t u = (uint32_t)(T[sl] >> 32);
      float f;
      std::memcpy(&f, &u, 4);
      th = (double)f;
    }
    put_int(ord_id[k]);
    s.push_back(' ');
    put_int((unsigned long long)v);
    s.push_back(' ');
    put_int((unsigned long long)c);
    s.push_back(' ');
    // std::to_chars with general format and precision 10 is printf's %.10g (the ostream precision(10) of
    // src/glue.cpp:49), several times faster; snprintf for the non-finite values a corrupt input could hold
    if (std::isfinite(th)) {
      auto r = std::to_chars(buf, buf + sizeof buf, th, std::chars_format::general, 10);
      s.append(buf, r.ptr);
    } else {
      const int len = std::snprintf(buf, sizeof buf, "%.10g", th);
      s.append(buf, buf + len);
    }
    s.push_back('\n');
  }
}
}  // namespace

extern "C" size_t fbh_tree_text(const uint64_t* trees, int ndpost, int m, int p, char* buf, size_t cap, int* varcount) {
  if (!trees || ndpost < 0 || m < 1 || p < 1) {
    set_error(FBH_EINVAL, "fbh_tree_text: bad argument");
    return 0;
  }
  std::string head = std::to_string(ndpost) + " " + std::to_string(m) + " " + std::to_string(p) + "\n";
  if (varcount) std::memset(varcount, 0, sizeof(int) * (size_t)ndpost * p);
  // draws are independent: format them in parallel, then concatenate in order
  unsigned nt = std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 16) nt = 16;
  if ((unsigned)ndpost < nt) nt = ndpost > 0 ? (unsigned)ndpost : 1;
  std::vector<std::string> part(nt);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t]() {
      const int lo = (int)((long)ndpost * t / nt), hi = (int)((long)ndpost * (t + 1) / nt);
      std::string& s = part[t];
      s.reserve((size_t)(hi - lo) * m * 64);
      for (int dr = lo; dr < hi; dr++)
        for (int j = 0; j < m; j++)
          tree_lines(trees + ((size_t)dr * m + j) * kNodes, p, s, varcount ? varcount + (size_t)dr * p : nullptr);
    });
  for (auto& x : th) x.join();
  size_t total = head.size();
  for (auto& s : part) total += s.size();
  if (buf && cap > 0) {
    size_t pos = 0;
    auto put = [&](const std::string& s) {
      const size_t k = std::min(s.size(), cap - pos);
      std::memcpy(buf + pos, s.data(), k);
      pos += k;
    };
    put(head);
    for (auto& s : part)
      if (pos < cap) put(s);
    if (total < cap) buf[total] = 0;
  }
  return total;
}

// ==================================================================== compile_trees / predict
struct fbh_trees {
  int device = 0;
  int ndpost = 0, ntree = 0, p = 0;
  // flat breadth-first arrays of src/predict.cpp:12-45 (absolute node indexes)
  std::vector<int> var, child, root, draw_lo;
  std::vector<double> cut, theta;
  // rank form: node words (next | var << 32 | c << 48, indexes relative to the draw), per-tree (root | depth << 24),
  // the cutpoint table the ranks are taken against; rank_ok when the table and the trees allow the rank form
  std::vector<uint64_t> node_w;
  std::vector<uint32_t> tinfo;
  std::vector<double> xi;
  std::vector<int> xoff;
  int pcut = 0, max_draw_nodes = 0;
  int max_var = -1;  // largest split variable (-1: no internal node)
  bool rank_ok = false;
  int *d_var = nullptr, *d_child = nullptr, *d_root = nullptr, *d_lo = nullptr;
  double *d_cut = nullptr, *d_theta = nullptr;
  uint64_t* d_node = nullptr;
  uint32_t* d_tinfo = nullptr;
  double* d_xi = nullptr;
  int* d_xoff = nullptr;
  // per-handle workspace; calls on one handle are serialised by mu and run in order on the caller's stream
  std::mutex mu;
  int* d_draws = nullptr;
  size_t draws_cap = 0;
  double* d_out = nullptr;
  size_t out_cap = 0;
  uint16_t* d_rank = nullptr;
  size_t rank_cap = 0;
  uint64_t rank_rows_uid = 0;  // the rows the ranks in d_rank belong to (0 = none)
  // Ordering of d_rank across streams: rank_ready is recorded after the k_rank launch that filled it (on rank_stream),
  // last_use after every predict kernel that read it (on last_stream). A launch on another stream waits for
  // rank_ready before reading the ranks; a refill waits for last_use before overwriting them.
  hipStream_t rank_stream = nullptr, last_stream = nullptr;
  hipEvent_t rank_ready = nullptr, last_use = nullptr;
  bool rank_recorded = false, use_recorded = false;
  // the kernel the last launch took: 1 rank form, 0 general form (-1: none yet); rows per thread of the rank form
  int last_form = -1, last_rows_per_thread = 0;
  // launch timing: a ring of event pairs, folded into the totals when a pair is reused
  static constexpr int kEv = 128;
  hipEvent_t ev0[kEv] = {}, ev1[kEv] = {};
  bool ev_live[kEv] = {};
  int ev_next = 0;
  double ev_ms = 0.0;
  long ev_count = 0;
};

struct fbh_rows {
  int device = 0, n = 0, p = 0;
  uint64_t uid = 0;
  double* d_x = nullptr;
};

static std::atomic<uint64_t> g_rows_uid{0};

static void free_trees(fbh_trees* t) {
  if (!t) return;
  (void)hipSetDevice(t->device);
  void* ptrs[] = {t->d_var, t->d_child, t->d_root, t->d_lo,  t->d_cut,   t->d_theta, t->d_node,
                  t->d_tinfo, t->d_xi,  t->d_xoff, t->d_draws, t->d_out, t->d_rank};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < fbh_trees::kEv; k++) {
    if (t->ev0[k]) (void)hipEventDestroy(t->ev0[k]);
    if (t->ev1[k]) (void)hipEventDestroy(t->ev1[k]);
  }
  if (t->rank_ready) (void)hipEventDestroy(t->rank_ready);
  if (t->last_use) (void)hipEventDestroy(t->last_use);
  delete t;
}

namespace {
struct Parser {
  const char* s;
  const char* e;
  bool ok = true;
  void ws() {
    while (s < e && (*s == ' ' || *s == '\n' || *s == '\t' || *s == '\r')) s++;
  }
  long long integer() {
    ws();
    long long v = 0;
    auto r = std::from_chars(s, e, v);
    if (r.ec != std::errc()) {
      ok = false;
      return 0;
    }
    s = r.ptr;
    return v;
  }
  double real() {
    ws();
    // strtod is correctly rounded (the value an istream >> double reads); copy the token to get a terminator
    const char* b = s;
    while (s < e && !(*s == ' ' || *s == '\n' || *s == '\t' || *s == '\r')) s++;
    std::string tok(b, s);
    char* end = nullptr;
    const double v = std::strtod(tok.c_str(), &end);
    if (tok.empty() || end != tok.c_str() + tok.size()) ok = false;
    return v;
  }
};
}  // namespace

// The rank form of compiled trees (see k_predict_rank): possible when every variable's cutpoints are nondecreasing
// and free of NaN with at most 65,535 of them, every tree is at most 255 levels deep and every draw holds fewer than
// 2^24 nodes. Otherwise predictions take the general kernel.
static void build_rank_form(fbh_trees* t, const double* cutpoints, const int* cut_off, int pcut,
                            const std::vector<int>& cidx, const std::vector<int>& tdepth) {
  t->rank_ok = false;
  t->pcut = pcut;
  t->xoff.assign(cut_off, cut_off + pcut + 1);
  t->xi.assign(cutpoints, cutpoints + (pcut > 0 ? cut_off[pcut] : 0));
  if (pcut > 65535) return;
  for (int v = 0; v < pcut; v++) {
    const int k0 = cut_off[v], k1 = cut_off[v + 1];
    if (k1 - k0 > 65535) return;
    for (int k = k0; k < k1; k++) {
      if (!(cutpoints[k] == cutpoints[k])) return;                 // NaN
      if (k > k0 && !(cutpoints[k - 1] <= cutpoints[k])) return;   // decreasing
    }
  }
  for (int d : tdepth)
    if (d > 255) return;
  t->max_draw_nodes = 0;
  for (int dr = 0; dr < t->ndpost; dr++) {
    const int cnt = t->draw_lo[dr + 1] - t->draw_lo[dr];
    if (cnt >= (1 << 24)) return;
    t->max_draw_nodes = std::max(t->max_draw_nodes, cnt);
  }
  const size_t nn = t->var.size();
  t->node_w.resize(nn);
  for (int dr = 0; dr < t->ndpost; dr++)
    for (int k = t->draw_lo[dr]; k < t->draw_lo[dr + 1]; k++) {
      const bool leaf = t->child[k] < 0;
      const uint64_t next = (uint64_t)(uint32_t)((leaf ? k : t->child[k]) - t->draw_lo[dr]);
      const uint64_t var = leaf ? 0u : (uint64_t)t->var[k], c = leaf ? kLeafCut : (uint64_t)cidx[k];
      t->node_w[k] = next | (var << 32) | (c << 48);
    }
  const long ntr = (long)t->ndpost * t->ntree;
  t->tinfo.resize(ntr);
  for (long tr = 0; tr < ntr; tr++)
    t->tinfo[tr] = (uint32_t)(t->root[tr] - t->draw_lo[tr / t->ntree]) | ((uint32_t)tdepth[tr] << 24);
  t->rank_ok = true;
}

static int upload_trees(fbh_trees* t);

extern "C" int fbh_compile_trees(const char* text, size_t len, const double* cutpoints, const int* cut_off, int pcut,
                                 int device, fbh_trees** out) {
  if (!text || !out || (pcut > 0 && (!cutpoints || !cut_off))) {
    set_error(FBH_EINVAL, "fbh_compile_trees: null argument");
    return FBH_EINVAL;
  }
  *out = nullptr;
  fbh_trees* t = new fbh_trees();
  t->device = device;
  Parser ps{text, text + len};
  t->ndpost = (int)ps.integer();
  t->ntree = (int)ps.integer();
  t->p = (int)ps.integer();
  if (!ps.ok || t->ndpost < 0 || t->ntree < 0) {
    set_error(FBH_EPARSE, "fbh_compile_trees: bad header");
    delete t;
    return FBH_EPARSE;
  }
  const long ntr = (long)t->ndpost * t->ntree;
  t->root.resize(ntr);
  t->draw_lo.assign(t->ndpost + 1, 0);
  struct Nd {
    unsigned long long id;
    int v, c;
    double th;
  };
  std::vector<Nd> nd;
  std::unordered_map<unsigned long long, int> pos;
  std::vector<int> qk, slot, cidx, depth;
  std::vector<int> tdepth(ntr, 0);
  for (long tr = 0; tr < ntr; tr++) {
    if (tr % t->ntree == 0) t->draw_lo[tr / t->ntree] = (int)t->var.size();
    const long long nn = ps.integer();
    if (!ps.ok || nn < 1 || nn > 1000000) {
      set_error(FBH_EPARSE, "fbh_compile_trees: bad node count");
      free_trees(t);
      return FBH_EPARSE;
    }
    nd.resize(nn);
    pos.clear();
    for (long long k = 0; k < nn; k++) {
      nd[k].id = (unsigned long long)ps.integer();
      nd[k].v = (int)ps.integer();
      nd[k].c = (int)ps.integer();
      nd[k].th = ps.real();
      pos[nd[k].id] = (int)k;
    }
    if (!ps.ok || !pos.count(1)) {
      set_error(FBH_EPARSE, "fbh_compile_trees: bad node line or no root");
      free_trees(t);
      return FBH_EPARSE;
    }
    // breadth first from id 1; the two children of a node are stored next to each other (src/predict.cpp:87-102)
    qk.assign(1, pos[1]);
    slot.assign(1, (int)t->var.size());
    t->root[tr] = slot[0];
    t->var.push_back(0);
    t->child.push_back(-1);
    t->cut.push_back(0);
    t->theta.push_back(0);
    cidx.push_back(0);
    depth.push_back(0);
    for (size_t h = 0; h < qk.size(); h++) {
      const int k = qk[h], me = slot[h];
      auto L = pos.find(2 * nd[k].id);
      t->theta[me] = nd[k].th;
      if (L != pos.end()) {
        auto R = pos.find(2 * nd[k].id + 1);
        const int v = nd[k].v, c = nd[k].c;
        if (R == pos.end() || v < 0 || v >= pcut || c < 0 || c >= cut_off[v + 1] - cut_off[v] ||
            qk.size() + 2 > (size_t)nn) {
          set_error(FBH_EPARSE, "fbh_compile_trees: inconsistent tree (missing child or cutpoint out of range)");
          free_trees(t);
          return FBH_EPARSE;
        }
        const int ch = (int)t->var.size();
        t->var[me] = v;
        t->cut[me] = cutpoints[cut_off[v] + c];
        t->child[me] = ch;
        cidx[me] = c;
        t->max_var = std::max(t->max_var, v);
        for (int z = 0; z < 2; z++) {
          t->var.push_back(0);
          t->child.push_back(-1);
          t->cut.push_back(0);
          t->theta.push_back(0);
          cidx.push_back(0);
          depth.push_back(depth[me] + 1);
        }
        tdepth[tr] = std::max(tdepth[tr], depth[me] + 1);
        qk.push_back(L->second);
        slot.push_back(ch);
        qk.push_back(R->second);
        slot.push_back(ch + 1);
      }
    }
  }
  t->draw_lo[t->ndpost] = (int)t->var.size();
  build_rank_form(t, cutpoints, cut_off, pcut, cidx, tdepth);
  // The device copies are made now when the device is visible, so that every prediction entry point (the
  // allocation-free one included) finds them in place; on a host without the device they wait for the first
  // prediction.
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) == hipSuccess && device >= 0 && device < ndev) {
    const int rc = upload_trees(t);
    if (rc) {
      free_trees(t);
      return rc;
    }
  }
  *out = t;
  return FBH_OK;
}

// device copies of the flat arrays (and of the rank form), made on the first prediction (parsing needs no device)
static int upload_trees(fbh_trees* t) {
  if (t->d_var) return FBH_OK;
  int rc = ensure_device(t->device);
  if (rc) return rc;
  hipError_t e = hipSuccess;
  auto up = [&](void** dst, const void* src, size_t bytes) {
    if (e == hipSuccess) e = hipMalloc(dst, bytes ? bytes : 8);
    if (e == hipSuccess && bytes) e = hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice);
  };
  up((void**)&t->d_child, t->child.data(), sizeof(int) * t->child.size());
  up((void**)&t->d_cut, t->cut.data(), sizeof(double) * t->cut.size());
  up((void**)&t->d_theta, t->theta.data(), sizeof(double) * t->theta.size());
  up((void**)&t->d_root, t->root.data(), sizeof(int) * t->root.size());
  up((void**)&t->d_lo, t->draw_lo.data(), sizeof(int) * t->draw_lo.size());
  if (t->rank_ok) {
    up((void**)&t->d_node, t->node_w.data(), sizeof(uint64_t) * t->node_w.size());
    up((void**)&t->d_tinfo, t->tinfo.data(), sizeof(uint32_t) * t->tinfo.size());
    up((void**)&t->d_xi, t->xi.data(), sizeof(double) * t->xi.size());
    up((void**)&t->d_xoff, t->xoff.data(), sizeof(int) * t->xoff.size());
  }
  up((void**)&t->d_var, t->var.data(), sizeof(int) * t->var.size());
  if (e != hipSuccess) {
    set_error(FBH_EHIP, std::string("fbh_predict: upload: ") + hipGetErrorString(e));
    return FBH_EHIP;
  }
  return FBH_OK;
}

extern "C" int fbh_trees_info(const fbh_trees* t, int* ndpost, int* ntree, int* p, int* nnodes) {
  if (!t) return FBH_EINVAL;
  if (ndpost) *ndpost = t->ndpost;
  if (ntree) *ntree = t->ntree;
  if (p) *p = t->p;
  if (nnodes) *nnodes = (int)t->var.size();
  return FBH_OK;
}

extern "C" int fbh_trees_arrays(const fbh_trees* t, int* var, int* child, double* cut, double* theta, int* root) {
  if (!t) return FBH_EINVAL;
  if (var) std::memcpy(var, t->var.data(), sizeof(int) * t->var.size());
  if (child) std::memcpy(child, t->child.data(), sizeof(int) * t->child.size());
  if (cut) std::memcpy(cut, t->cut.data(), sizeof(double) * t->cut.size());
  if (theta) std::memcpy(theta, t->theta.data(), sizeof(double) * t->theta.size());
  if (root) std::memcpy(root, t->root.data(), sizeof(int) * t->root.size());
  return FBH_OK;
}

extern "C" void fbh_trees_free(fbh_trees* t) { free_trees(t); }

extern "C" int fbh_rows_upload(const double* const* cols, int n, int p, int device, fbh_rows** out) {
  if (!out || n < 0 || p < 0 || (p > 0 && !cols)) {
    set_error(FBH_EINVAL, "fbh_rows_upload: bad argument");
    return FBH_EINVAL;
  }
  int rc = ensure_device(device);
  if (rc) return rc;
  fbh_rows* r = new fbh_rows();
  r->uid = ++g_rows_uid;
  r->device = device;
  r->n = n;
  r->p = p;
  if (hipMalloc(&r->d_x, sizeof(double) * ((size_t)n * p ? (size_t)n * p : 1)) != hipSuccess) {
    delete r;
    set_error(FBH_ENOMEM, "fbh_rows_upload: hipMalloc failed");
    return FBH_ENOMEM;
  }
  for (int v = 0; v < p; v++)
    if (n && hipMemcpy(r->d_x + (size_t)v * n, cols[v], sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipFree(r->d_x);
      delete r;
      set_error(FBH_EHIP, "fbh_rows_upload: copy failed");
      return FBH_EHIP;
    }
  *out = r;
  return FBH_OK;
}

extern "C" void fbh_rows_free(fbh_rows* r) {
  if (!r) return;
  (void)hipSetDevice(r->device);
  (void)hipFree(r->d_x);
  delete r;
}

// argument checks shared by the predict entry points (draws on the host)
static int check_predict(const fbh_trees* t, const fbh_rows* r, int ndraws) {
  if (!t || !r || ndraws < 0) {
    set_error(FBH_EINVAL, "fbh_predict: bad argument");
    return FBH_EINVAL;
  }
  if (r->p != t->p) {
    set_error(FBH_EINVAL, "fbh_predict: newdata has " + std::to_string(r->p) + " columns, the trees use " +
                              std::to_string(t->p));
    return FBH_EINVAL;
  }
  if (r->device != t->device) {
    set_error(FBH_EINVAL, "fbh_predict: rows and trees live on different devices");
    return FBH_EINVAL;
  }
  if (t->max_var >= r->p) {
    set_error(FBH_EINVAL, "fbh_predict: a split variable is outside the newdata columns");
    return FBH_EINVAL;
  }
  return FBH_OK;
}

static bool grow(void** buf, size_t* cap, size_t bytes) {
  if (*cap >= bytes) return true;
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  if (hipMalloc(buf, bytes) != hipSuccess) return false;
  *cap = bytes;
  return true;
}

// LDS bytes of one k_predict_rank workgroup for draws of up to `nodes` nodes
static size_t rank_lds_bytes(int nodes, int ntree, int p, int rows) {
  return (size_t)nodes * 16 + (size_t)((ntree + 3) & ~3) * 4 + (size_t)p * rows * 2;
}
// Dynamic LDS the rank kernel may use on a device: the device's per-workgroup limit, raised for the kernel once per
// device (0 if that fails, which sends every prediction to the general kernel).
static size_t rank_lds_limit(int device) {
  static std::mutex m;
  static size_t lim[64];
  static bool done[64];
  if (device < 0 || device >= 64) return 0;
  std::lock_guard<std::mutex> lk(m);
  if (!done[device]) {
    done[device] = true;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(k_predict_rank<1>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, v) == hipSuccess &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(k_predict_rank<2>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, v) == hipSuccess)
      lim[device] = (size_t)v;
  }
  return lim[device];
}

// The predict kernel of ndraws draws (device array d_draws, 1-based) for the rows r into d_out [ndraws][n] on stream
// s: the rank form reads the rows' ranks from rank (filled by k_rank), the general form the rows themselves. Caller
// holds t->mu.
static int predict_kernels(fbh_trees* t, const fbh_rows* r, const int* d_draws, int ndraws, double mu, double* d_out,
                           hipStream_t s, bool rank_form, int rows_per_thread, size_t lds, const uint16_t* rank) {
  t->last_form = rank_form ? 1 : 0;
  t->last_rows_per_thread = rank_form ? rows_per_thread : 1;
  if (rank_form) {
    const int rows = kRankThreads * rows_per_thread;
    dim3 grid((r->n + rows - 1) / rows, ndraws);
    if (rows_per_thread == 2)
      hipLaunchKernelGGL(k_predict_rank<2>, grid, dim3(kRankThreads), lds, s, t->d_node, t->d_theta, t->d_tinfo,
                         t->d_lo, d_draws, rank, r->n, r->p, t->ntree, t->ndpost, mu, d_out);
    else
      hipLaunchKernelGGL(k_predict_rank<1>, grid, dim3(kRankThreads), lds, s, t->d_node, t->d_theta, t->d_tinfo,
                         t->d_lo, d_draws, rank, r->n, r->p, t->ntree, t->ndpost, mu, d_out);
  } else {
    dim3 grid((r->n + kPredThreads - 1) / kPredThreads, ndraws);
    hipLaunchKernelGGL(k_predict, grid, dim3(kPredThreads), 0, s, t->d_var, t->d_child, t->d_cut, t->d_theta,
                       t->d_root, d_draws, t->d_lo, r->d_x, r->n, t->ntree, t->ndpost, mu, d_out);
  }
  FBH_HIP_CHECK(hipGetLastError());
  return FBH_OK;
}

// Launch the prediction of ndraws draws (device array d_draws, 1-based, checked by the caller) for the rows r into
// d_out [ndraws][n] on stream s. Caller holds t->mu.
static int predict_launch(fbh_trees* t, const fbh_rows* r, const int* d_draws, int ndraws, double mu, double* d_out,
                          hipStream_t s) {
  if (ndraws == 0 || r->n == 0) return FBH_OK;
  int rc = upload_trees(t);
  if (rc) return rc;
  FBH_HIP_CHECK(hipSetDevice(t->device));
  // fold the oldest event pair of the ring into the totals before reusing it
  const int slot = t->ev_next;
  t->ev_next = (t->ev_next + 1) % fbh_trees::kEv;
  if (!t->ev0[slot]) {
    FBH_HIP_CHECK(hipEventCreate(&t->ev0[slot]));
    FBH_HIP_CHECK(hipEventCreate(&t->ev1[slot]));
  }
  if (t->ev_live[slot]) {
    float ms = 0;
    FBH_HIP_CHECK(hipEventSynchronize(t->ev1[slot]));
    if (hipEventElapsedTime(&ms, t->ev0[slot], t->ev1[slot]) == hipSuccess) {
      t->ev_ms += ms;
      t->ev_count++;
    }
    t->ev_live[slot] = false;
  }
  // two rows per thread when that still gives every CU two workgroups, else one (few draws: more, shorter blocks)
  const long blocks2 = (long)((r->n + kRankRowsMax - 1) / kRankRowsMax) * ndraws;
  const int rows_per_thread = blocks2 >= 512 ? 2 : 1;
  const size_t lds = rank_lds_bytes(t->max_draw_nodes, t->ntree, r->p, kRankThreads * rows_per_thread);
  const size_t lds_max = rank_lds_limit(t->device);
  const bool rank_form = t->rank_ok && lds <= lds_max;
  if (rank_form && !t->rank_ready) {
    FBH_HIP_CHECK(hipEventCreateWithFlags(&t->rank_ready, hipEventDisableTiming));
    FBH_HIP_CHECK(hipEventCreateWithFlags(&t->last_use, hipEventDisableTiming));
  }
  if (rank_form && t->rank_rows_uid != r->uid) {
    // the rows' ranks against this table; rows are immutable, so the ranks stay valid while the same rows are used.
    // A launch still reading the old ranks on another stream finishes first (a larger buffer is a new allocation,
    // and hipFree of the old one waits for the device).
    if (t->use_recorded && t->last_stream != s) FBH_HIP_CHECK(hipStreamWaitEvent(s, t->last_use, 0));
    if (!grow((void**)&t->d_rank, &t->rank_cap, sizeof(uint16_t) * (size_t)r->n * (r->p ? r->p : 1))) {
      set_error(FBH_ENOMEM, "fbh_predict: hipMalloc of the rank buffer failed");
      return FBH_ENOMEM;
    }
    if (r->p > 0) {
      hipLaunchKernelGGL(k_rank, dim3((r->n + 255) / 256, r->p), dim3(256), 0, s, r->d_x, t->d_xi, t->d_xoff, r->n,
                         r->p, t->pcut, t->d_rank);
      FBH_HIP_CHECK(hipGetLastError());
    }
    FBH_HIP_CHECK(hipEventRecord(t->rank_ready, s));
    t->rank_recorded = true;
    t->rank_stream = s;
    t->rank_rows_uid = r->uid;
  } else if (rank_form && t->rank_recorded && t->rank_stream != s) {
    // ranks computed on another stream: this launch reads them only once they are complete
    FBH_HIP_CHECK(hipStreamWaitEvent(s, t->rank_ready, 0));
  }
  FBH_HIP_CHECK(hipEventRecord(t->ev0[slot], s));
  rc = predict_kernels(t, r, d_draws, ndraws, mu, d_out, s, rank_form, rows_per_thread, lds, t->d_rank);
  if (rc) return rc;
  FBH_HIP_CHECK(hipEventRecord(t->ev1[slot], s));
  t->ev_live[slot] = true;
  if (rank_form) {
    FBH_HIP_CHECK(hipEventRecord(t->last_use, s));
    t->use_recorded = true;
    t->last_stream = s;
  }
  return FBH_OK;
}

extern "C" int fbh_predict_last_launch(const fbh_trees* tc, int* form, int* rows_per_thread) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  if (!t) {
    set_error(FBH_EINVAL, "fbh_predict_last_launch: null trees");
    return FBH_EINVAL;
  }
  std::lock_guard<std::mutex> lk(t->mu);
  if (form) *form = t->last_form;
  if (rows_per_thread) *rows_per_thread = t->last_rows_per_thread;
  return FBH_OK;
}

extern "C" int fbh_predict_rows(const fbh_trees* tc, const fbh_rows* r, const int* draws, int ndraws, double mu,
                                double* out_host) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  int rc = check_predict(t, r, ndraws);
  if (rc) return rc;
  if (!draws || !out_host) {
    set_error(FBH_EINVAL, "fbh_predict: null draws or output");
    return FBH_EINVAL;
  }
  for (int k = 0; k < ndraws; k++)
    if (draws[k] < 1 || draws[k] > t->ndpost) {
      set_error(FBH_EINVAL, "fbh_predict: draw " + std::to_string(draws[k]) + " out of range 1.." +
                                std::to_string(t->ndpost));
      return FBH_EINVAL;
    }
  if (ndraws == 0 || r->n == 0) return FBH_OK;
  std::lock_guard<std::mutex> lk(t->mu);
  rc = ensure_device(t->device);
  if (rc) return rc;
  FBH_HIP_CHECK(hipSetDevice(t->device));
  if (!grow((void**)&t->d_draws, &t->draws_cap, sizeof(int) * ndraws) ||
      !grow((void**)&t->d_out, &t->out_cap, sizeof(double) * (size_t)ndraws * r->n)) {
    set_error(FBH_ENOMEM, "fbh_predict: hipMalloc of the workspace failed");
    return FBH_ENOMEM;
  }
  FBH_HIP_CHECK(hipMemcpy(t->d_draws, draws, sizeof(int) * ndraws, hipMemcpyHostToDevice));
  rc = predict_launch(t, r, t->d_draws, ndraws, mu, t->d_out, nullptr);
  if (rc) return rc;
  FBH_HIP_CHECK(hipStreamSynchronize(nullptr));
  FBH_HIP_CHECK(hipMemcpy(out_host, t->d_out, sizeof(double) * (size_t)ndraws * r->n, hipMemcpyDeviceToHost));
  return FBH_OK;
}

extern "C" int fbh_predict_rows_device(const fbh_trees* tc, const fbh_rows* r, const int* draws_dev, int ndraws,
                                       double mu, double* out_dev, void* stream) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  int rc = check_predict(t, r, ndraws);
  if (rc) return rc;
  if (ndraws > 0 && (!draws_dev || !out_dev)) {
    set_error(FBH_EINVAL, "fbh_predict_rows_device: null draws or output");
    return FBH_EINVAL;
  }
  std::lock_guard<std::mutex> lk(t->mu);
  return predict_launch(t, r, draws_dev, ndraws, mu, out_dev, reinterpret_cast<hipStream_t>(stream));
}

// ------------------------------------------------------------------ allocation-free prediction
// The rows' ranks of fbh_predict_rows_device_noalloc live in `scratch`, a file-scope device array: every GPU the library
// runs on holds its own copy (at its own address) until the process ends. Host code never passes the array's name to a
// kernel (on the host it is only a stand-in symbol): the device address comes from hipGetSymbolAddress, taken once per
// device with that device current, and only that address is used.
constexpr size_t kNoAllocRanks = size_t(8) << 20;  // 8 Mi ranks = 16 MiB: rows x predictors of one call at most
static __device__ __align__(256) uint16_t scratch[kNoAllocRanks];
static std::mutex g_scratch_mu;
static uint16_t* g_scratch_addr[64];  // per device ordinal: the device address of its copy of scratch

static int scratch_address(int device, uint16_t** out) {
  if (device < 0 || device >= 64) {
    set_error(FBH_EINVAL, "fbh_predict_rows_device_noalloc: device ordinal out of range");
    return FBH_EINVAL;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    FBH_HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    FBH_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
      set_error(FBH_EHIP, "fbh_predict_rows_device_noalloc: no device address for the rank scratch");
      return FBH_EHIP;
    }
    g_scratch_addr[device] = static_cast<uint16_t*>(p);
  }
  *out = g_scratch_addr[device];
  return FBH_OK;
}

extern "C" int fbh_predict_rows_device_noalloc(const fbh_trees* tc, const fbh_rows* r, const int* draws_dev,
                                               int ndraws, double mu, double* out_dev, void* stream) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  int rc = check_predict(t, r, ndraws);
  if (rc) return rc;
  if (ndraws > 0 && (!draws_dev || !out_dev)) {
    set_error(FBH_EINVAL, "fbh_predict_rows_device_noalloc: null draws or output");
    return FBH_EINVAL;
  }
  if ((size_t)r->n * (size_t)r->p > kNoAllocRanks) {
    set_error(FBH_EINVAL, "fbh_predict_rows_device_noalloc: rows x predictors = " +
                              std::to_string((size_t)r->n * (size_t)r->p) + " exceeds the scratch of " +
                              std::to_string(kNoAllocRanks) + " ranks");
    return FBH_EINVAL;
  }
  if (ndraws == 0 || r->n == 0) return FBH_OK;
  std::lock_guard<std::mutex> lk(t->mu);
  if (!t->d_var) {
    set_error(FBH_EINVAL, "fbh_predict_rows_device_noalloc: the trees are not on the device (compiled while no "
                          "device was visible)");
    return FBH_EINVAL;
  }
  uint16_t* ranks = nullptr;
  if ((rc = scratch_address(t->device, &ranks))) return rc;
  FBH_HIP_CHECK(hipSetDevice(t->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long blocks2 = (long)((r->n + kRankRowsMax - 1) / kRankRowsMax) * ndraws;
  const int rows_per_thread = blocks2 >= 512 ? 2 : 1;
  const size_t lds = rank_lds_bytes(t->max_draw_nodes, t->ntree, r->p, kRankThreads * rows_per_thread);
  const bool rank_form = t->rank_ok && lds <= rank_lds_limit(t->device);
  if (rank_form && r->p > 0) {
    // every rank the predict kernel reads is written here first: nothing carries over from an earlier call
    hipLaunchKernelGGL(k_rank, dim3((r->n + 255) / 256, r->p), dim3(256), 0, s, r->d_x, t->d_xi, t->d_xoff, r->n, r->p,
                       t->pcut, ranks);
    FBH_HIP_CHECK(hipGetLastError());
  }
  return predict_kernels(t, r, draws_dev, ndraws, mu, out_dev, s, rank_form, rows_per_thread, lds, ranks);
}

extern "C" int fbh_predict_timing(const fbh_trees* tc, double* avg_ms, int* nlaunch) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  if (!t) return FBH_EINVAL;
  std::lock_guard<std::mutex> lk(t->mu);
  double tot = t->ev_ms;
  long k = t->ev_count;
  for (int s = 0; s < fbh_trees::kEv; s++) {
    if (!t->ev_live[s]) continue;
    float ms = 0;
    if (hipEventSynchronize(t->ev1[s]) == hipSuccess && hipEventElapsedTime(&ms, t->ev0[s], t->ev1[s]) == hipSuccess) {
      tot += ms;
      k++;
    }
  }
  if (avg_ms) *avg_ms = k ? tot / k : 0;
  if (nlaunch) *nlaunch = (int)k;
  return FBH_OK;
}

extern "C" int fbh_predict_draws(const fbh_trees* t, const double* const* cols, int n, int p, const int* draws,
                                 int ndraws, double mu, double* out) {
  if (!t) {
    set_error(FBH_EINVAL, "fbh_predict_draws: null trees");
    return FBH_EINVAL;
  }
  fbh_rows* r = nullptr;
  int rc = fbh_rows_upload(cols, n, p, t->device, &r);
  if (rc) return rc;
  rc = fbh_predict_rows(t, r, draws, ndraws, mu, out);
  fbh_rows_free(r);
  return rc;
}

extern "C" int fbh_predict_draw(const fbh_trees* t, const double* const* cols, int n, int p, int m_draw, double mu,
                                double* out) {
  return fbh_predict_draws(t, cols, n, p, &m_draw, 1, mu, out);
}

extern "C" int fbh_trees_rank_form(const fbh_trees* t) { return t && t->rank_ok ? 1 : 0; }

extern "C" int fbh_device_alloc(size_t bytes, int device, void** out) {
  if (!out) return FBH_EINVAL;
  *out = nullptr;
  int rc = ensure_device(device);
  if (rc) return rc;
  FBH_HIP_CHECK(hipSetDevice(device));
  if (hipMalloc(out, bytes ? bytes : 8) != hipSuccess) {
    set_error(FBH_ENOMEM, "fbh_device_alloc: hipMalloc failed");
    return FBH_ENOMEM;
  }
  return FBH_OK;
}

extern "C" void fbh_device_free(void* p, int device) {
  if (!p) return;
  (void)hipSetDevice(device);
  (void)hipFree(p);
}

extern "C" int fbh_device_copy(void* dst, const void* src, size_t bytes, int kind, int device) {
  if ((!dst || !src) && bytes) return FBH_EINVAL;
  FBH_HIP_CHECK(hipSetDevice(device));
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                       : hipMemcpyDeviceToDevice;
  FBH_HIP_CHECK(hipMemcpy(dst, src, bytes, k));
  return FBH_OK;
}

extern "C" int fbh_device_sync(int device) {
  FBH_HIP_CHECK(hipSetDevice(device));
  FBH_HIP_CHECK(hipDeviceSynchronize());
  return FBH_OK;
}
