// Host check of the FFT building blocks of tinker-gpu_amd/csrc/hip/fft_core.hpp:
// the Stockham stages with the compile-time plans of every grid length the
// device kernels specialize (radix 8 first, then 4, 2, 3, 5), both transform
// directions, against a naive DFT.  Built and run by tests/test_fft_host.py.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>
#include "hip/fft_core.hpp"
using namespace t9;
template <int N>
double check(int sign) {
  const int nl = 3, P = N + 1;
  std::vector<Cplx<double>> a(nl * P), b(nl * P), tw(N);
  for (int m = 0; m < N; ++m) tw[m] = {std::cos(-2 * M_PI * m / N), std::sin(-2 * M_PI * m / N)};
  std::mt19937 g(5);
  std::normal_distribution<double> d;
  std::vector<std::complex<double>> x(nl * N);
  for (int l = 0; l < nl; ++l)
    for (int e = 0; e < N; ++e) {
      x[l * N + e] = {d(g), d(g)};
      a[l * P + e] = {x[l * N + e].real(), x[l * N + e].imag()};
    }
  FftPlan plan{};
  plan.n = N;
  plan.nstage = ct_nstage(N);
  for (int st = 0; st < plan.nstage; ++st) plan.radix[st] = ct_radix(N, st);
  struct NoBarrier { void operator()() const {} };
  Cplx<double>* r = fft_stages<double>(a.data(), b.data(), nl, P, plan, tw.data(), sign, 0, 1, NoBarrier());
  double err = 0;
  for (int l = 0; l < nl; ++l)
    for (int q = 0; q < N; ++q) {
      std::complex<double> s = 0;
      for (int e = 0; e < N; ++e) s += x[l * N + e] * std::polar(1.0, sign * 2 * M_PI * e * q / N);
      err = std::fmax(err, std::abs(s - std::complex<double>(r[l * P + q].x, r[l * P + q].y)));
    }
  return err;
}
int main() {
  double e = 0;
  for (int s : {-1, 1}) {
    e = std::fmax(e, check<24>(s)); e = std::fmax(e, check<32>(s)); e = std::fmax(e, check<40>(s));
    e = std::fmax(e, check<48>(s)); e = std::fmax(e, check<64>(s)); e = std::fmax(e, check<80>(s));
    e = std::fmax(e, check<96>(s)); e = std::fmax(e, check<120>(s)); e = std::fmax(e, check<128>(s));
    e = std::fmax(e, check<160>(s));
  }
  std::printf("max error %.3e\n", e);
  return e < 1e-10 ? 0 : 1;
}
