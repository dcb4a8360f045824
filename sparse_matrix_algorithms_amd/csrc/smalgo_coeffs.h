// Host-side filter coefficients (see smalgo_coeffs.cpp for reference lines).
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace smalgo {

std::vector<double> jackson_kernel(int K);
std::vector<double> cheb_delta_coeffs(double sigma_hat, int K, bool jackson);
std::vector<double> cheb_window_coeffs(double a, double b, int K, bool jackson);
std::vector<double> cheb_smooth_window_coeffs(double a, double b, int K, bool jackson, int nsamp,
                                              double delta);
std::vector<double> build_coeffs(int degree, double smin, double smax, double sigma, double emin,
                                 double emax, int window_type, bool jackson);

}  // namespace smalgo
