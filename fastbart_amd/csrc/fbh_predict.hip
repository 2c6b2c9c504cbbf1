// Output formats and prediction: treedraws text and varcount (src/glue.cpp:29-51), per-draw training fits
// (src/glue.cpp:52-56) on the device, and compile_trees / predict_draw (src/predict.cpp) with a tree-outer device
// kernel over flat child-pair arrays.
#include <charconv>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fbh_internal.h"

namespace fbh {

// ------------------------------------------------------------------ training fits from packed words
// One thread per (draw, row); the m trees of the draw are added in order to a double that starts at 0.0, each tree
// walked with q <= c going left. The draw's words are read through the cache (every thread of the draw reads the same
// words).
__global__ __launch_bounds__(256) void k_yhat(const uint64_t* __restrict__ trees, const uint8_t* __restrict__ q, int n,
                                              int m, double* __restrict__ out) {
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
      const int s = q[(size_t)v * n + i] <= c ? (int)((w >> 16) & 255) : (int)((w >> 24) & 255);
      w = T[s];
    }
    f += (double)__uint_as_float((uint32_t)(w >> 32));
  }
  out[(size_t)dr * n + i] = f;
}

// ------------------------------------------------------------------ predictor kernel
// Tree-outer over a block of rows: the workgroup stages the flat node arrays of one draw in LDS (when they fit) and
// every thread adds the draw's trees in order for its row; x >= cut goes to the right child (src/predict.cpp:125-133).
constexpr int kPredThreads = 256;
constexpr int kPredLdsNodes = 2048;  // 2048 nodes x (int var, int child, double cut, double theta) = 48 KiB

__global__ __launch_bounds__(kPredThreads) void k_predict(const int* __restrict__ var, const int* __restrict__ child,
                                                          const double* __restrict__ cut,
                                                          const double* __restrict__ theta,
                                                          const int* __restrict__ root, const int* __restrict__ draws,
                                                          const int* __restrict__ draw_lo, const double* __restrict__ x,
                                                          int n, int ntree, double mu, double* __restrict__ out) {
  __shared__ int s_var[kPredLdsNodes];
  __shared__ int s_child[kPredLdsNodes];
  __shared__ double s_cut[kPredLdsNodes];
  __shared__ double s_theta[kPredLdsNodes];
  const int k = blockIdx.y;
  const int dr = draws[k] - 1;
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
int yhat_device(const uint64_t* d_trees, const uint8_t* d_q, int n, int ndpost, int m, hipStream_t stream,
                double* host_out) {
  if (ndpost == 0 || n == 0) return FBH_OK;
  double* d_out = nullptr;
  FBH_HIP_CHECK(hipMalloc(&d_out, sizeof(double) * (size_t)ndpost * n));
  dim3 grid((n + 255) / 256, ndpost);
  hipLaunchKernelGGL(k_yhat, grid, dim3(256), 0, stream, d_trees, d_q, n, m, d_out);
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
  rc = yhat_device(dt, dq, n, ndpost, m, nullptr, yhat);
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
    const int len = std::snprintf(buf, sizeof buf, "%.10g", th);
    s.append(buf, buf + len);
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
  std::vector<int> var, child, root, draw_lo;
  std::vector<double> cut, theta;
  int *d_var = nullptr, *d_child = nullptr, *d_root = nullptr, *d_lo = nullptr;
  double *d_cut = nullptr, *d_theta = nullptr;
  std::vector<hipEvent_t> ev;
};

struct fbh_rows {
  int device = 0, n = 0, p = 0;
  double* d_x = nullptr;
};

static void free_trees(fbh_trees* t) {
  if (!t) return;
  (void)hipSetDevice(t->device);
  void* ptrs[] = {t->d_var, t->d_child, t->d_root, t->d_lo, t->d_cut, t->d_theta};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto e : t->ev) (void)hipEventDestroy(e);
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
  std::vector<int> qk, slot;
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
        for (int z = 0; z < 2; z++) {
          t->var.push_back(0);
          t->child.push_back(-1);
          t->cut.push_back(0);
          t->theta.push_back(0);
        }
        qk.push_back(L->second);
        slot.push_back(ch);
        qk.push_back(R->second);
        slot.push_back(ch + 1);
      }
    }
  }
  t->draw_lo[t->ndpost] = (int)t->var.size();
  *out = t;
  return FBH_OK;
}

// device copies of the flat arrays, made on the first prediction (parsing needs no device)
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

extern "C" int fbh_predict_rows(const fbh_trees* tc, const fbh_rows* r, const int* draws, int ndraws, double mu,
                                double* out_host) {
  fbh_trees* t = const_cast<fbh_trees*>(tc);
  if (!t || !r || !draws || ndraws < 0 || !out_host) {
    set_error(FBH_EINVAL, "fbh_predict: bad argument");
    return FBH_EINVAL;
  }
  if (r->p != t->p) {
    set_error(FBH_EINVAL, "fbh_predict: newdata has " + std::to_string(r->p) + " columns, the trees use " +
                              std::to_string(t->p));
    return FBH_EINVAL;
  }
  for (int k = 0; k < ndraws; k++)
    if (draws[k] < 1 || draws[k] > t->ndpost) {
      set_error(FBH_EINVAL, "fbh_predict: draw " + std::to_string(draws[k]) + " out of range 1.." +
                                std::to_string(t->ndpost));
      return FBH_EINVAL;
    }
  for (size_t k = 0; k < t->var.size(); k++)
    if (t->child[k] >= 0 && (t->var[k] < 0 || t->var[k] >= r->p)) {
      set_error(FBH_EINVAL, "fbh_predict: a split variable is outside the newdata columns");
      return FBH_EINVAL;
    }
  if (ndraws == 0 || r->n == 0) return FBH_OK;
  int rc = upload_trees(t);
  if (rc) return rc;
  FBH_HIP_CHECK(hipSetDevice(t->device));
  int* d_draws = nullptr;
  double* d_out = nullptr;
  FBH_HIP_CHECK(hipMalloc(&d_draws, sizeof(int) * ndraws));
  FBH_HIP_CHECK(hipMalloc(&d_out, sizeof(double) * (size_t)ndraws * r->n));
  FBH_HIP_CHECK(hipMemcpy(d_draws, draws, sizeof(int) * ndraws, hipMemcpyHostToDevice));
  hipEvent_t e0, e1;
  FBH_HIP_CHECK(hipEventCreate(&e0));
  FBH_HIP_CHECK(hipEventCreate(&e1));
  t->ev.push_back(e0);
  t->ev.push_back(e1);
  dim3 grid((r->n + kPredThreads - 1) / kPredThreads, ndraws);
  FBH_HIP_CHECK(hipEventRecord(e0, nullptr));
  hipLaunchKernelGGL(k_predict, grid, dim3(kPredThreads), 0, nullptr, t->d_var, t->d_child, t->d_cut, t->d_theta,
                     t->d_root, d_draws, t->d_lo, r->d_x, r->n, t->ntree, mu, d_out);
  FBH_HIP_CHECK(hipGetLastError());
  FBH_HIP_CHECK(hipEventRecord(e1, nullptr));
  FBH_HIP_CHECK(hipDeviceSynchronize());
  FBH_HIP_CHECK(hipMemcpy(out_host, d_out, sizeof(double) * (size_t)ndraws * r->n, hipMemcpyDeviceToHost));
  (void)hipFree(d_draws);
  (void)hipFree(d_out);
  return FBH_OK;
}

extern "C" int fbh_predict_timing(const fbh_trees* t, double* avg_ms, int* nlaunch) {
  if (!t) return FBH_EINVAL;
  double tot = 0;
  int k = 0;
  for (size_t i = 0; i + 1 < t->ev.size(); i += 2) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t->ev[i], t->ev[i + 1]) == hipSuccess) {
      tot += ms;
      k++;
    }
  }
  if (avg_ms) *avg_ms = k ? tot / k : 0;
  if (nlaunch) *nlaunch = k;
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
