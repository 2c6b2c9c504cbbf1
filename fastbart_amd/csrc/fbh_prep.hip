// Preprocessing of the R wrappers in C++ (host code only): fwbart's R/fwbart.R:8-27 and fpbart's R/fpbart.R:7-21, so
// that callers without R (the batched C-ABI, services) can hand the library raw x / y / w.
//
//   model matrix (BART::bartModelMatrix defaults): constant columns dropped (rm.const), per-column cutpoints -- the
//     midpoints between consecutive distinct values when a column has at most numcut + 1 of them, else numcut points
//     equally spaced strictly inside [min, max]; caller cutpoints (xinfo) replace them
//   q = findInterval(x, xi): the number of cutpoints <= x (R/fwbart.R:19)
//   fwbart: fmean = mean(y); yc = y - fmean; sigest = the residual standard error of lm(yc ~ X) when p < n, else
//     sd(yc); lambda = sigest^2 qchisq(1 - sigquant, nu) / nu; tau = (max yc - min yc) / (2 k sqrt(ntree))
//     (R/fwbart.R:21-28)
//   fpbart: offset = qnorm(mean(y)); tau = 3 / (k sqrt(ntree)) (R/fpbart.R:7, :17)
//   pgtab = base / (1 + d)^power, d = 0..63; the sampler takes y, w, pgtab and the scalars as float (src/glue.cpp:16,27)
//
// Arithmetic follows R's where R fixes it: mean() and sd() accumulate in long double with mean's second (correction)
// pass; x^2 is x * x; lm's rank comes from a Householder QR that, like R's LINPACK dqrdc2, keeps the column order and
// treats a column as aliased when its norm after the previous columns' reflections falls below 1e-7 of its original
// norm (residual degrees of freedom n - rank). qnorm is Wichura's AS 241 (PPND16); qchisq inverts the regularised
// incomplete gamma function (series / continued fraction, DLMF 8.7.1 and 8.9.2) by safeguarded Newton steps from the
// Wilson-Hilferty start.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "fbh_internal.h"

namespace fbh {
namespace {

// mean of R's summary.c: a long double sum, then the mean of the residuals added back
double r_mean(const double* x, int n) {
  long double s = 0.0L;
  for (int i = 0; i < n; i++) s += x[i];
  s /= n;
  long double t = 0.0L;
  for (int i = 0; i < n; i++) t += x[i] - s;
  s += t / n;
  return (double)s;
}

// sd(): the square root of the two-pass variance about the mean, accumulated in long double
double r_sd(const double* x, int n) {
  const double m = r_mean(x, n);
  long double ss = 0.0L;
  for (int i = 0; i < n; i++) {
    const double d = x[i] - m;  // a double product added to a long double sum, as cov.c's
    ss += d * d;
  }
  return std::sqrt((double)(ss / (n - 1)));
}

// Wichura, M. J. (1988) Algorithm AS 241: The percentage points of the normal distribution. Applied Statistics 37,
// 477-484 (PPND16, about 16 significant digits).
double qnorm_as241(double p) {
  if (!(p > 0.0 && p < 1.0)) {
    if (p == 0.0) return -std::numeric_limits<double>::infinity();
    if (p == 1.0) return std::numeric_limits<double>::infinity();
    return std::numeric_limits<double>::quiet_NaN();
  }
  const double q = p - 0.5;
  double r, val;
  if (std::fabs(q) <= 0.425) {
    r = 0.180625 - q * q;
    val = q * (((((((r * 2509.0809287301226727 + 33430.575583588128105) * r + 67265.770927008700853) * r +
                    45921.953931549871457) * r + 13731.693765509461125) * r + 1971.5909503065514427) * r +
                 133.14166789178437745) * r + 3.387132872796366608) /
          (((((((r * 5226.495278852545925 + 28729.085735721942674) * r + 39307.89580009271061) * r +
               21213.794301586595867) * r + 5394.1960214247511077) * r + 687.1870074920579083) * r +
            42.313330701600911252) * r + 1.0);
    return val;
  }
  r = q < 0 ? p : 1.0 - p;
  r = std::sqrt(-std::log(r));
  if (r <= 5.0) {
    r -= 1.6;
    val = (((((((r * 7.7454501427834140764e-4 + 0.0227238449892691845833) * r + 0.24178072517745061177) * r +
                1.27045825245236838258) * r + 3.64784832476320460504) * r + 5.7694972214606914055) * r +
             4.6303378461565452959) * r + 1.42343711074968357734) /
          (((((((r * 1.05075007164441684324e-9 + 5.475938084995344946e-4) * r + 0.0151986665636164571966) * r +
               0.14810397642748007459) * r + 0.68976733498510000455) * r + 1.6763848301838038494) * r +
            2.05319162663775882187) * r + 1.0);
  } else {
    r -= 5.0;
    val = (((((((r * 2.01033439929228813265e-7 + 2.71155556874348757815e-5) * r + 0.0012426609473880784386) * r +
                0.026532189526576123093) * r + 0.29656057182850489123) * r + 1.7848265399172913358) * r +
             5.4637849111641143699) * r + 6.6579046435011037772) /
          (((((((r * 2.04426310338993978564e-15 + 1.4215117583164458887e-7) * r + 1.8463183175100546818e-5) * r +
               7.868691311456132591e-4) * r + 0.0148753612908506148525) * r + 0.13692988092273580531) * r +
            0.59983220655588793769) * r + 1.0);
  }
  return q < 0.0 ? -val : val;
}

// log of the integrand's prefactor x^a e^-x / Gamma(a)
double log_prefactor(double a, double x) { return a * std::log(x) - x - std::lgamma(a); }

// regularised incomplete gamma functions P(a, x) (lower) and Q(a, x) = 1 - P (upper): the power series of P (DLMF
// 8.7.1) below x = a + 1, else the continued fraction of Q (DLMF 8.9.2, modified Lentz); the other one by difference,
// which is accurate there as it is at least about 0.4
void pgamma_pq(double a, double x, double* P, double* Q) {
  if (x <= 0.0) {
    *P = 0.0;
    *Q = 1.0;
    return;
  }
  if (x < a + 1.0) {
    double term = 1.0 / a, sum = term;
    for (int k = 1; k < 10000; k++) {
      term *= x / (a + k);
      sum += term;
      if (std::fabs(term) < std::fabs(sum) * 1e-17) break;
    }
    *P = sum * std::exp(log_prefactor(a, x));
    *Q = 1.0 - *P;
    return;
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int k = 1; k < 10000; k++) {
    const double an = -k * (k - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-17) break;
  }
  *Q = std::exp(log_prefactor(a, x)) * h;
  *P = 1.0 - *Q;
}

// qchisq(p, df): the x with P(df / 2, x / 2) = p, by Newton steps kept inside a shrinking bracket; for p above 1/2 the
// equation is solved as Q(df / 2, x / 2) = 1 - p (exact in double there), so that the upper tail keeps its digits
double qchisq_host(double p, double df) {
  if (!(df > 0.0) || !(p >= 0.0 && p <= 1.0)) return std::numeric_limits<double>::quiet_NaN();
  if (p == 0.0) return 0.0;
  if (p == 1.0) return std::numeric_limits<double>::infinity();
  const double a = 0.5 * df;
  // Wilson-Hilferty start; for small p the leading term of the series, P(a, y) ~ y^a / Gamma(a + 1)
  const double z = qnorm_as241(p), h = 2.0 / (9.0 * df);
  double x = df * std::pow(1.0 - h + z * std::sqrt(h), 3.0);
  if (!(x > 0.0) || (p < 0.05 && df < 10.0)) {
    const double xs = 2.0 * std::exp((std::log(p) + std::lgamma(a + 1.0)) / a);
    if (!(x > 0.0) || xs < x) x = xs;
  }
  const bool upper = p > 0.5;
  const double pu = 1.0 - p;
  double lo = 0.0, hi = std::numeric_limits<double>::infinity();
  for (int it = 0; it < 200; it++) {
    double P, Q;
    pgamma_pq(a, 0.5 * x, &P, &Q);
    const double f = upper ? pu - Q : P - p;  // increasing in x either way
    if (f == 0.0) return x;
    if (f < 0.0) lo = std::max(lo, x);
    else hi = std::min(hi, x);
    // d/dx P(a, x / 2) = (x / 2)^(a - 1) e^(-x / 2) / (2 Gamma(a))
    const double dens = std::exp((a - 1.0) * std::log(0.5 * x) - 0.5 * x - std::lgamma(a)) * 0.5;
    double xn = x - f / dens;
    if (!(xn > lo && xn < hi)) xn = std::isfinite(hi) ? 0.5 * (lo + hi) : 2.0 * std::max(x, lo);
    xn = std::min(std::max(xn, 0.25 * x), 4.0 * x);  // at most a factor of 4 per step
    if (std::fabs(xn - x) <= 4.0 * std::numeric_limits<double>::epsilon() * xn) return xn;
    x = xn;
  }
  return x;
}

// Residual standard error of lm(y ~ 1 + X): Householder QR of [1, X] (n x (p + 1)) in column order, a column aliased
// when its remaining norm is below tol times its original norm; sqrt(RSS / (n - rank)).
double lm_sigma(const std::vector<const double*>& cols, int n, const double* y, int* rank_out) {
  const int k = (int)cols.size() + 1;
  std::vector<double> A((size_t)n * k), b(y, y + n);
  for (int i = 0; i < n; i++) A[i] = 1.0;
  for (int j = 1; j < k; j++) std::memcpy(&A[(size_t)j * n], cols[j - 1], sizeof(double) * n);
  std::vector<double> norm0(k);
  for (int j = 0; j < k; j++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += A[(size_t)j * n + i] * A[(size_t)j * n + i];
    norm0[j] = std::sqrt(s);
  }
  const double tol = 1e-7;
  int r = 0;  // reflections applied = rank so far; they act on rows r.. of the remaining columns and of y
  for (int j = 0; j < k && r < n; j++) {
    double* c = &A[(size_t)j * n];
    double s = 0.0;
    for (int i = r; i < n; i++) s += c[i] * c[i];
    const double nrm = std::sqrt(s);
    if (!(nrm > tol * norm0[j]) || nrm == 0.0) continue;  // aliased with the columns before it
    // reflector v = c[r..] + sign(c[r]) |c| e_r, H = I - 2 v v^T / (v^T v)
    const double alpha = c[r] >= 0.0 ? -nrm : nrm;
    std::vector<double> v(c + r, c + n);
    v[0] -= alpha;
    double vv = 0.0;
    for (double e : v) vv += e * e;
    if (vv == 0.0) {
      r++;
      continue;
    }
    auto reflect = [&](double* z) {
      double d = 0.0;
      for (int i = r; i < n; i++) d += v[i - r] * z[i];
      const double f = 2.0 * d / vv;
      for (int i = r; i < n; i++) z[i] -= f * v[i - r];
    };
    for (int jj = j; jj < k; jj++) reflect(&A[(size_t)jj * n]);
    reflect(b.data());
    r++;
  }
  if (rank_out) *rank_out = r;
  double rss = 0.0;
  for (int i = r; i < n; i++) rss += b[i] * b[i];
  return n > r ? std::sqrt(rss / (n - r)) : std::numeric_limits<double>::quiet_NaN();
}

struct Prep {
  fbh_inputs out{};  // handed to the caller; out.priv_ points back here
  std::vector<uint8_t> q;
  std::vector<int> numcut, rmc, cut_off, ybin;
  std::vector<double> cuts;
  std::vector<float> y, wt;
};

// the model matrix: kept columns, their cutpoints and cutpoint ranks
int model_matrix(const double* const* x, int n, int ncol, const double* xinfo, const int* xinfo_off, int numcut,
                 int rm_const, Prep& P, std::vector<const double*>& kept) {
  P.cut_off.assign(1, 0);
  for (int v = 0; v < ncol; v++) {
    const double* col = x[v];
    for (int i = 0; i < n; i++)
      if (!std::isfinite(col[i])) {
        set_error(FBH_EINVAL, "fbh_*_inputs: non-finite x in column " + std::to_string(v + 1));
        return FBH_EINVAL;
      }
    std::vector<double> u(col, col + n);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    if (rm_const && u.size() < 2) continue;
    std::vector<double> xi;
    if (xinfo) {
      xi.assign(xinfo + xinfo_off[v], xinfo + xinfo_off[v + 1]);
      for (size_t t = 0; t < xi.size(); t++)
        if (!std::isfinite(xi[t]) || (t > 0 && xi[t] < xi[t - 1])) {
          set_error(FBH_EINVAL, "fbh_*_inputs: xinfo of column " + std::to_string(v + 1) +
                                    " is not a finite nondecreasing sequence (findInterval needs one)");
          return FBH_EINVAL;
        }
    } else if (u.size() >= 2) {
      if ((int)u.size() <= numcut + 1) {
        for (size_t t = 0; t + 1 < u.size(); t++) xi.push_back((u[t] + u[t + 1]) / 2.0);
      } else {
        // seq(min, max, length.out = numcut + 2) without its end points: min + t by, by = (max - min) / (numcut + 1)
        const double by = (u.back() - u.front()) / (numcut + 1);
        for (int t = 1; t <= numcut; t++) xi.push_back(u.front() + t * by);
      }
    }
    if (xi.size() > 255) {
      set_error(FBH_EINVAL, "fbh_*_inputs: column " + std::to_string(v + 1) + " has " + std::to_string(xi.size()) +
                                " cutpoints (at most 255; R falls back to BART here)");
      return FBH_EINVAL;
    }
    const size_t pk = P.rmc.size();
    P.q.resize((pk + 1) * (size_t)n);
    for (int i = 0; i < n; i++)
      P.q[pk * n + i] = (uint8_t)(std::upper_bound(xi.begin(), xi.end(), col[i]) - xi.begin());
    P.cuts.insert(P.cuts.end(), xi.begin(), xi.end());
    P.cut_off.push_back((int)P.cuts.size());
    P.numcut.push_back((int)xi.size());
    P.rmc.push_back(v + 1);
    kept.push_back(col);
  }
  if (P.rmc.empty() || (int)P.rmc.size() > FBH_PMAX) {
    set_error(FBH_EINVAL, "fbh_*_inputs: " + std::to_string(P.rmc.size()) + " predictors after rm.const (1.." +
                              std::to_string(FBH_PMAX) + " supported; R falls back to BART here)");
    return FBH_EINVAL;
  }
  return FBH_OK;
}

int check_common(const double* const* x, int n, int ncol, const double* xinfo, const int* xinfo_off,
                 const fbh_prep_opts* o, fbh_inputs** out) {
  if (!out || !o || !x || n < 2 || ncol < 1 || (xinfo && !xinfo_off)) {
    set_error(FBH_EINVAL, "fbh_*_inputs: null argument or fewer than 2 rows");
    return FBH_EINVAL;
  }
  *out = nullptr;
  if (n > FBH_NMAX || o->ntree < 1 || o->ntree > FBH_MMAX || o->numcut < 1 || !(o->k > 0.0)) {
    set_error(FBH_EINVAL, "fbh_*_inputs: unsupported settings (n <= 65536, 1 <= ntree <= 200, numcut >= 1, k > 0; R "
                          "falls back to BART here)");
    return FBH_EINVAL;
  }
  return FBH_OK;
}

void finish(Prep* P, int n, const fbh_prep_opts* o) {
  fbh_inputs& r = P->out;
  r.n = n;
  r.p = (int)P->rmc.size();
  r.q = P->q.data();
  r.numcut = P->numcut.data();
  r.rm_const = P->rmc.data();
  r.cutpoints = P->cuts.data();
  r.cut_off = P->cut_off.data();
  r.y = P->y.data();
  r.wt = P->wt.data();
  r.ybin = P->ybin.empty() ? nullptr : P->ybin.data();
  for (int d = 0; d < FBH_NODES; d++) {
    const double b = 1.0 + d;
    const double den = o->power == 2.0 ? b * b : std::pow(b, o->power);
    r.pgtab[d] = (float)(o->base / den);
  }
  r.tau = (float)r.tau_d;
  r.lambda = (float)r.lambda_d;
  r.nu = (float)r.nu_d;
  r.sigest = (float)r.sigest_d;
  r.offset = (float)r.offset_d;
}

}  // namespace
}  // namespace fbh

using namespace fbh;

extern "C" void fbh_prep_defaults(int probit, fbh_prep_opts* o) {
  if (!o) return;
  const double na = std::numeric_limits<double>::quiet_NaN();
  o->ntree = probit ? 50 : 200;
  o->numcut = 100;
  o->rm_const = 1;
  o->k = 2.0;
  o->power = 2.0;
  o->base = 0.95;
  o->sigdf = 3.0;
  o->sigquant = 0.90;
  o->sigest = na;
  o->sigmaf = na;
  o->lambda = na;
  o->fmean = na;
  o->binary_offset = na;
}

extern "C" int fbh_wbart_inputs(const double* const* x, int n, int ncol, const double* y, const double* w,
                                const double* xinfo, const int* xinfo_off, const fbh_prep_opts* o, fbh_inputs** out) {
  int rc = check_common(x, n, ncol, xinfo, xinfo_off, o, out);
  if (rc) return rc;
  if (!y) {
    set_error(FBH_EINVAL, "fbh_wbart_inputs: null y");
    return FBH_EINVAL;
  }
  for (int i = 0; i < n; i++)
    if (!std::isfinite(y[i]) || (w && !(std::isfinite(w[i]) && w[i] != 0.0))) {
      set_error(FBH_EINVAL, "fbh_wbart_inputs: non-finite y or w (or w = 0) at row " + std::to_string(i + 1));
      return FBH_EINVAL;
    }
  Prep* P = new Prep();
  std::vector<const double*> kept;
  if ((rc = model_matrix(x, n, ncol, xinfo, xinfo_off, o->numcut, o->rm_const, *P, kept))) {
    delete P;
    return rc;
  }
  const int p = (int)kept.size();
  fbh_inputs& r = P->out;
  r.fmean = std::isnan(o->fmean) ? r_mean(y, n) : o->fmean;
  std::vector<double> yc(n);
  double ymax = -std::numeric_limits<double>::infinity(), ymin = std::numeric_limits<double>::infinity();
  for (int i = 0; i < n; i++) {
    yc[i] = y[i] - r.fmean;
    ymax = std::max(ymax, yc[i]);
    ymin = std::min(ymin, yc[i]);
  }
  r.nu_d = o->sigdf;
  if (std::isnan(o->lambda)) {
    double sg = o->sigest;
    if (std::isnan(sg)) sg = p < n ? lm_sigma(kept, n, yc.data(), &r.lm_rank) : r_sd(yc.data(), n);
    r.sigest_d = sg;
    r.lambda_d = (sg * sg * qchisq_host(1.0 - o->sigquant, r.nu_d)) / r.nu_d;
  } else {
    r.lambda_d = o->lambda;
    r.sigest_d = std::sqrt(o->lambda);
  }
  r.tau_d = std::isnan(o->sigmaf) ? (ymax - ymin) / (2.0 * o->k * std::sqrt((double)o->ntree))
                                  : o->sigmaf / std::sqrt((double)o->ntree);
  r.offset_d = 0.0;
  P->y.resize(n);
  P->wt.resize(n);
  for (int i = 0; i < n; i++) {
    P->y[i] = (float)yc[i];
    const double wi = w ? w[i] : 1.0;
    P->wt[i] = (float)(1.0 / (wi * wi));
  }
  finish(P, n, o);
  P->out.priv_ = P;
  *out = &P->out;
  return FBH_OK;
}

extern "C" int fbh_pbart_inputs(const double* const* x, int n, int ncol, const int* y01, const double* xinfo,
                                const int* xinfo_off, const fbh_prep_opts* o, fbh_inputs** out) {
  int rc = check_common(x, n, ncol, xinfo, xinfo_off, o, out);
  if (rc) return rc;
  if (!y01) {
    set_error(FBH_EINVAL, "fbh_pbart_inputs: null y");
    return FBH_EINVAL;
  }
  long long ones = 0;
  for (int i = 0; i < n; i++) {
    if (y01[i] != 0 && y01[i] != 1) {
      set_error(FBH_EINVAL, "fbh_pbart_inputs: y is not 0/1 at row " + std::to_string(i + 1));
      return FBH_EINVAL;
    }
    ones += y01[i];
  }
  Prep* P = new Prep();
  std::vector<const double*> kept;
  if ((rc = model_matrix(x, n, ncol, xinfo, xinfo_off, o->numcut, o->rm_const, *P, kept))) {
    delete P;
    return rc;
  }
  fbh_inputs& r = P->out;
  // mean of an integer vector: a long double sum over n (summary.c), then qnorm
  r.offset_d = std::isnan(o->binary_offset) ? qnorm_as241((double)((long double)ones / n)) : o->binary_offset;
  r.fmean = 0.0;
  r.tau_d = 3.0 / (o->k * std::sqrt((double)o->ntree));
  r.lambda_d = 1.0;
  r.nu_d = 3.0;
  r.sigest_d = 1.0;
  P->y.assign(n, 0.f);
  P->wt.assign(n, 1.f);
  P->ybin.assign(y01, y01 + n);
  finish(P, n, o);
  P->out.priv_ = P;
  *out = &P->out;
  return FBH_OK;
}

extern "C" void fbh_inputs_free(fbh_inputs* in) {
  if (!in) return;
  delete static_cast<Prep*>(in->priv_);
}

extern "C" double fbh_qnorm(double p) { return qnorm_as241(p); }
extern "C" double fbh_qchisq(double p, double df) { return qchisq_host(p, df); }
