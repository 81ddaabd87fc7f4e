// Host-side substitution-model setup for the engine.
//
// Mirrors what raxml-ng / libpll prepare before any CLV update
// (src/RaxmlWrapper.cpp:189-214, 615-645 set the model on the
// pllmod_treeinfo): a reversible rate matrix normalised to mean rate 1, its
// eigendecomposition (P = V exp(L r t) V^-1, as libpll's
// pll_update_prob_matrices uses), and equiprobable discrete Gamma
// categories with mean rates (PLL_GAMMA_RATES_MEAN, RaxmlWrapper.cpp:621).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace nrx {

namespace {

// Regularised lower incomplete gamma P(a, x).
double inc_gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double lg = std::lgamma(a);
  if (x < a + 1.0) {
    double term = 1.0 / a, sum = term;
    for (int n = 1; n < 200000; ++n) {
      term *= x / (a + n);
      sum += term;
      if (std::fabs(term) <= std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(a * std::log(x) - x - lg);
  }
  // continued fraction for Q(a, x) (modified Lentz)
  const double tiny = 1e-300;
  double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, f = d;
  for (int i = 1; i < 200000; ++i) {
    const double an = -static_cast<double>(i) * (i - a);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double delta = c * d;
    f *= delta;
    if (std::fabs(delta - 1.0) <= 1e-17) break;
  }
  return 1.0 - std::exp(a * std::log(x) - x - lg) * f;
}

// Quantile of Gamma(shape a, scale 1): Wilson-Hilferty start, safeguarded
// Newton iterations on P(a, x) - p with a bracketing fallback.
double gamma_quantile(double a, double p) {
  // normal quantile by a rational approximation refined with Newton on erfc
  auto norm_q = [](double pp) {
    double t = std::sqrt(-2.0 * std::log(pp < 0.5 ? pp : 1.0 - pp));
    double z = t - (2.515517 + 0.802853 * t + 0.010328 * t * t) /
                       (1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t);
    z = pp < 0.5 ? -z : z;
    for (int it = 0; it < 3; ++it) {
      const double cdf = 0.5 * std::erfc(-z / std::sqrt(2.0));
      const double pdf = std::exp(-0.5 * z * z) / std::sqrt(2.0 * M_PI);
      z -= (cdf - pp) / pdf;
    }
    return z;
  };
  const double z = norm_q(p);
  double x = a * std::pow(1.0 - 1.0 / (9.0 * a) + z / (3.0 * std::sqrt(a)), 3);
  if (!(x > 0.0)) x = std::pow(p * std::tgamma(a + 1.0), 1.0 / a);
  double lo = 0.0, hi = std::max(2.0 * x, a + 50.0);
  while (inc_gamma_p(a, hi) < p) hi *= 2.0;
  const double lg = std::lgamma(a);
  for (int it = 0; it < 200; ++it) {
    const double f = inc_gamma_p(a, x) - p;
    if (f < 0.0)
      lo = x;
    else
      hi = x;
    const double dens = std::exp((a - 1.0) * std::log(x) - x - lg);
    double nx = x - f / dens;
    if (!(nx > lo && nx < hi)) nx = 0.5 * (lo + hi);
    if (std::fabs(nx - x) <= 1e-16 * std::max(1.0, x)) {
      x = nx;
      break;
    }
    x = nx;
  }
  return x;
}

}  // namespace

std::vector<double> gamma_mean_rates(double alpha, int ncat) {
  std::vector<double> r(ncat, 1.0);
  if (ncat == 1) return r;
  if (!(alpha > 0.0)) throw std::invalid_argument("Gamma alpha must be positive");
  double prev = 0.0;
  for (int k = 0; k < ncat; ++k) {
    double cur = 1.0;
    if (k + 1 < ncat) {
      // quantile of Gamma(alpha, rate alpha) at (k+1)/ncat, times alpha
      const double xa = gamma_quantile(alpha, static_cast<double>(k + 1) / ncat);
      cur = inc_gamma_p(alpha + 1.0, xa);
    }
    r[k] = (cur - prev) * ncat;
    prev = cur;
  }
  return r;
}

// Cyclic Jacobi eigendecomposition of a symmetric S x S matrix (row-major).
// On return A is destroyed, evals holds eigenvalues, U the eigenvectors as
// columns.
static void jacobi_eigen(int S, std::vector<double> &A, std::vector<double> &evals,
                         std::vector<double> &U) {
  U.assign(S * S, 0.0);
  for (int i = 0; i < S; ++i) U[i * S + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < S; ++i)
      for (int j = i + 1; j < S; ++j) off += A[i * S + j] * A[i * S + j];
    if (off < 1e-300) break;
    for (int p = 0; p < S; ++p)
      for (int q = p + 1; q < S; ++q) {
        const double apq = A[p * S + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = A[p * S + p], aqq = A[q * S + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < S; ++k) {
          const double akp = A[k * S + p], akq = A[k * S + q];
          A[k * S + p] = c * akp - s * akq;
          A[k * S + q] = s * akp + c * akq;
        }
        for (int k = 0; k < S; ++k) {
          const double apk = A[p * S + k], aqk = A[q * S + k];
          A[p * S + k] = c * apk - s * aqk;
          A[q * S + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < S; ++k) {
          const double ukp = U[k * S + p], ukq = U[k * S + q];
          U[k * S + p] = c * ukp - s * ukq;
          U[k * S + q] = s * ukp + c * ukq;
        }
      }
  }
  evals.resize(S);
  for (int i = 0; i < S; ++i) evals[i] = A[i * S + i];
}

void build_model(int S, const double *subst, const double *freqs, EigenModel &m) {
  for (int i = 0; i < S; ++i)
    if (!(freqs[i] > 0.0))
      throw std::invalid_argument("state frequencies must be positive");
  const int nr = S * (S - 1) / 2;
  for (int k = 0; k < nr; ++k)
    if (!(subst[k] >= 0.0)) throw std::invalid_argument("substitution rates must be >= 0");
  // Q_ij = s_ij pi_j, normalised so that -sum_i pi_i Q_ii = 1
  std::vector<double> Q(S * S, 0.0);
  int k = 0;
  for (int i = 0; i < S; ++i)
    for (int j = i + 1; j < S; ++j, ++k) {
      Q[i * S + j] = subst[k] * freqs[j];
      Q[j * S + i] = subst[k] * freqs[i];
    }
  double mu = 0.0;
  for (int i = 0; i < S; ++i) {
    double row = 0.0;
    for (int j = 0; j < S; ++j)
      if (j != i) row += Q[i * S + j];
    Q[i * S + i] = -row;
    mu += freqs[i] * row;
  }
  if (!(mu > 0.0)) throw std::invalid_argument("degenerate rate matrix");
  // symmetrised A = D^1/2 Q D^-1/2 with D = diag(pi)
  std::vector<double> A(S * S);
  for (int i = 0; i < S; ++i)
    for (int j = 0; j < S; ++j)
      A[i * S + j] = Q[i * S + j] / mu * std::sqrt(freqs[i]) / std::sqrt(freqs[j]);
  for (int i = 0; i < S; ++i)
    for (int j = i + 1; j < S; ++j) {
      const double avg = 0.5 * (A[i * S + j] + A[j * S + i]);
      A[i * S + j] = A[j * S + i] = avg;
    }
  std::vector<double> U, ev;
  jacobi_eigen(S, A, ev, U);
  m.S = S;
  m.eval = ev;
  m.V.assign(S * S, 0.0);
  m.Vinv.assign(S * S, 0.0);
  for (int i = 0; i < S; ++i)
    for (int kk = 0; kk < S; ++kk) {
      m.V[i * S + kk] = U[i * S + kk] / std::sqrt(freqs[i]);   // D^-1/2 U
      m.Vinv[kk * S + i] = U[i * S + kk] * std::sqrt(freqs[i]);  // U^T D^1/2
    }
  m.freqs.assign(freqs, freqs + S);
}

}  // namespace nrx
