// Host-side Chebyshev filter coefficients and spectral map.
//
// jackson_kernel      polfed.jl:275-284   (chebyshev_subspace_iteration.jl:376-385)
// cheb_delta_coeffs   polfed.jl:299-313   (chebyshev_subspace_iteration.jl:400-414)
// cheb_window_coeffs  polfed.jl:329-347   (chebyshev_subspace_iteration.jl:430-448)
// cheb_smooth_window  polfed.jl:364-388   (chebyshev_subspace_iteration.jl:465-489)
// _build_coeffs       polfed.jl:404-414
#include "smalgo_coeffs.h"
#include "smalgo_internal.h"

#include <algorithm>
#include <cmath>

namespace smalgo {

std::vector<double> jackson_kernel(int K) {
  if (K <= 0) return {1.0};
  std::vector<double> g(K + 1);
  const double a = M_PI / (K + 2);
  const double sa = std::sin(a);
  for (int k = 0; k <= K; ++k)
    g[k] = std::sin((k + 1) * a) / ((K + 2) * sa) + (1.0 - double(k + 1) / (K + 2)) * std::cos(k * a);
  return g;
}

static void damp(std::vector<double>& c, int K, bool jackson) {
  if (!jackson) return;
  const std::vector<double> g = jackson_kernel(K);
  for (int k = 0; k <= K; ++k) c[k] *= g[k];
}

std::vector<double> cheb_delta_coeffs(double sh, int K, bool jackson) {
  std::vector<double> c(std::max(K, 0) + 1, 0.0);
  const double th = std::acos(std::min(std::max(sh, -1.0), 1.0));
  c[0] = 1.0 / M_PI;
  for (int k = 1; k <= K; ++k) c[k] = 2.0 * std::cos(k * th) / M_PI;
  damp(c, K, jackson);
  return c;
}

std::vector<double> cheb_window_coeffs(double a, double b, int K, bool jackson) {
  double aa = std::max(-1.0, std::min(1.0, a));
  double bb = std::max(-1.0, std::min(1.0, b));
  if (aa > bb) std::swap(aa, bb);
  const double al = std::acos(bb), be = std::acos(aa);
  std::vector<double> c(std::max(K, 0) + 1, 0.0);
  c[0] = (be - al) / M_PI;
  for (int k = 1; k <= K; ++k) c[k] = 2.0 * (std::sin(k * be) - std::sin(k * al)) / (k * M_PI);
  damp(c, K, jackson);
  return c;
}

std::vector<double> cheb_smooth_window_coeffs(double a, double b, int K, bool jackson, int nsamp,
                                              double delta) {
  if (nsamp <= 1) nsamp = 4096;
  if (!(delta > 0)) delta = 0.02;
  double aa = std::max(-1.0, std::min(1.0, a));
  double bb = std::max(-1.0, std::min(1.0, b));
  if (aa > bb) std::swap(aa, bb);
  const double w = M_PI / (nsamp - 1);
  std::vector<double> f(nsamp), th(nsamp);
  for (int i = 0; i < nsamp; ++i) {
    th[i] = i * w;
    const double xi = std::cos(th[i]);
    const double wt = (i == 0 || i == nsamp - 1) ? 0.5 : 1.0;
    f[i] = wt * 0.5 * (1 + std::tanh((xi - aa) / delta)) * 0.5 * (1 - std::tanh((xi - bb) / delta));
  }
  std::vector<double> c(std::max(K, 0) + 1, 0.0);
  for (int k = 0; k <= K; ++k) {
    double s = 0.0;
    for (int i = 0; i < nsamp; ++i) s += f[i] * std::cos(k * th[i]);
    c[k] = (k == 0 ? 1.0 / M_PI : 2.0 / M_PI) * w * s;
  }
  damp(c, K, jackson);
  return c;
}

std::vector<double> build_coeffs(int degree, double smin, double smax, double sigma, double emin,
                                 double emax, int window_type, bool jackson) {
  const double c = 0.5 * (smin + smax);
  const double d = 0.5 * (smax - smin);
  const double sh = (sigma - c) / d;
  const double a = (emin - c) / d;
  const double b = (emax - c) / d;
  switch (window_type) {
    case 0: return cheb_window_coeffs(a, b, degree, jackson);
    case 1: return cheb_delta_coeffs(sh, degree, jackson);
    case 2: return cheb_smooth_window_coeffs(a, b, degree, jackson, 4096, 0.02);
    default: throw Error(-1, "Unknown window type: " + std::to_string(window_type));
  }
}

}  // namespace smalgo
