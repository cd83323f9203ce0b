// Stochastic-trace mode: the reference's default estimator for systems too
// large for the exact trace.
//   * runKPM (withSP2_deltaH.cpp:1562-1633): Chebyshev moments of the trace,
//     momentsTrace[m] = sum_r r^T T_m(H~) r over nRandoms Gaussian vectors;
//   * the DOS on the 2*Nmoments Chebyshev nodes with the Jackson kernel, the
//     chemical potential at the electron-count fraction of the DOS area and the
//     tightened spectral bounds (:1634-1712);
//   * computeDiagonalElements in random mode (:1296-1399):
//     diagP = sum_r r .* (P r) ./ sum_r r .* r.
// The vectors come from the SplitMix64 -> Box-Muller stream of a seed (the
// reference seeds std::mt19937 from std::random_device, so its draws are not
// reproducible; the estimator is the same).  The moments of all vectors are
// one block recursion on the device (the exact-trace kernels with explicit
// start vectors and moment doubling); P r is getKPE on the vector block.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "system.hpp"

namespace {

constexpr double kPi = 3.141592653589793238463;

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
void to_rowmajor(const double* R, int nvec, int norb, int ld, std::vector<T>& out) {
  out.assign((size_t)norb * ld, T(0));
  for (int cidx = 0; cidx < nvec; cidx++)
    for (int j = 0; j < norb; j++) out[(size_t)j * ld + cidx] = (T)R[(size_t)cidx * norb + j];
}

}  // namespace

extern "C" {

int ltb_kpe_block_moments(ltb_ctx* c, int M, int nvec, const double* R, int precision, double* moments,
                          double* trace) {
  if (!c || !R || M < 1 || nvec < 1 || (precision != LTB_FP64 && precision != LTB_FP32_ACC64))
    return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  const int norb = c->norb, ld = kpe_block_width(c, precision, nvec);
  const size_t esz = precision == LTB_FP64 ? 8 : 4;
  int st;
  if ((st = ensure_buffer(c, &c->d_xa, &c->xabytes, (size_t)norb * ld * esz))) return st;
  if ((st = ensure_buffer(c, &c->d_xb, &c->xbbytes, (size_t)norb * ld * esz))) return st;
  // start vectors row-major [orbital][column] in the recursion's precision, and
  // their squared norms in FP64 (the m = 0 moment and the even-moment offset)
  std::vector<double> r2(ld, 0.0);
  for (int k = 0; k < nvec; k++) {
    double s = 0;
    for (int j = 0; j < norb; j++) {
      const double x = precision == LTB_FP64 ? R[(size_t)k * norb + j] : (double)(float)R[(size_t)k * norb + j];
      s += x * x;
    }
    r2[k] = s;
  }
  if (precision == LTB_FP64) {
    std::vector<double> h;
    to_rowmajor(R, nvec, norb, ld, h);
    LTB_HIP(c, hipMemcpy(c->d_xa, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h;
    to_rowmajor(R, nvec, norb, ld, h);
    LTB_HIP(c, hipMemcpy(c->d_xa, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  DevBuf d_r2, d_mom;
  LTB_HIP(c, hipMalloc(&d_r2.p, sizeof(double) * ld));
  LTB_HIP(c, hipMalloc(&d_mom.p, sizeof(double) * (size_t)(M + 1) * ld));
  LTB_HIP(c, hipMemcpy(d_r2.p, r2.data(), sizeof(double) * ld, hipMemcpyHostToDevice));
  if ((st = kpe_block_moments_impl(c, M, nvec, precision, (const double*)d_r2.p, (double*)d_mom.p, ld))) return st;
  std::vector<double> mom((size_t)(M + 1) * ld);
  LTB_HIP(c, hipMemcpyAsync(mom.data(), d_mom.p, sizeof(double) * mom.size(), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  for (int m = 0; m <= M; m++) {
    double s = 0;
    for (int k = 0; k < nvec; k++) {
      const double v = mom[(size_t)m * ld + k];
      if (moments) moments[(size_t)k * (M + 1) + m] = v;
      s += v;
    }
    if (trace) trace[m] = s;
  }
  return LTB_OK;
}

int ltb_random_vectors(unsigned long long seed, int nvec, int n, double* R) {
  if (nvec < 1 || n < 1 || !R) return LTB_ERR_ARG;
  ltb::gaussian_fill(seed, (size_t)nvec * n, R);
  return LTB_OK;
}

int ltb_kpe_stochastic(ltb_ctx* c, int M, int nvec, unsigned long long seed, int precision, double* trace) {
  if (!c || nvec < 1) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  std::vector<double> R((size_t)nvec * c->norb);
  ltb::gaussian_fill(seed, R.size(), R.data());
  return ltb_kpe_block_moments(c, M, nvec, R.data(), precision, nullptr, trace);
}

int ltb_density_stochastic(ltb_ctx* c, int M, const double* w, int nvec, unsigned long long seed,
                           int precision, double* diagP) {
  if (!c || !w || !diagP || nvec < 1) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  const int norb = c->norb;
  std::vector<double> R((size_t)nvec * norb), Y((size_t)nvec * norb);
  ltb::gaussian_fill(seed, R.size(), R.data());
  int st = ltb_kpe_apply(c, M, w, nvec, R.data(), Y.data(), precision);
  if (st) return st;
  for (int j = 0; j < norb; j++) {
    double num = 0, den = 0;
    for (int k = 0; k < nvec; k++) {
      const double r = R[(size_t)k * norb + j];
      num += r * Y[(size_t)k * norb + j];
      den += r * r;
    }
    diagP[j] = num / den;
  }
  return LTB_OK;
}

int ltb_kpm_dos(const double* moments, int nmom, int ntilde, double a, double b, int nsize, double nelectrons,
                double* energies, double* dos, double* mu, double* emin, double* emax) {
  if (!moments || nmom < 2 || ntilde < 2 || !(a > 0) || nsize < 1) return LTB_ERR_ARG;
  // Jackson kernel with N = nmom (:1593-1595) and nodes x_k = cos(pi (k + 1/2) / ntilde) (:1588-1590)
  std::vector<double> g(nmom), x(ntilde), d(ntilde);
  const double t = kPi / (nmom + 1);
  for (int n = 0; n < nmom; n++)
    g[n] = ((nmom - n + 1) * std::cos(t * n) + std::sin(t * n) * std::cos(t) / std::sin(t)) / (nmom + 1);
  for (int k = 0; k < ntilde; k++) x[k] = std::cos(kPi * (k + 0.5) / ntilde);
  for (int k = 0; k < ntilde; k++) {
    // sum_j g_j mu_j T_j(x) by the Chebyshev recurrence (the reference evaluates cos(j acos x))
    double tm1 = 1.0, t0 = x[k], s = g[0] * moments[0] + 2 * g[1] * moments[1] * t0;
    for (int j = 2; j < nmom; j++) {
      const double tj = 2 * x[k] * t0 - tm1;
      s += 2 * g[j] * moments[j] * tj;
      tm1 = t0;
      t0 = tj;
    }
    d[k] = s / (kPi * std::sqrt(1.0 - x[k] * x[k]));
  }
  // area, fraction Ne/(2N) of it from the bottom of the band, bounds where
  // DOS/N >= 0.001 (:1660-1701)
  double all = 0;
  for (int i = ntilde - 1; i > 0; i--) all += (x[i - 1] - x[i]) * (d[i] + d[i - 1]) / 2;
  const double frac = nelectrons * all / (2.0 * nsize);
  double acc = 0;
  int imu = 0;
  for (int i = ntilde - 1; i > 0; i--) {
    acc += (x[i - 1] - x[i]) * (d[i] + d[i - 1]) / 2;
    if (acc >= frac) {
      imu = i;
      break;
    }
  }
  double lo = -10, hi = 10;
  for (int i = ntilde - 1; i > 0; i--)
    if (d[i] / nsize >= 0.001) {
      lo = x[i] * a + b;
      break;
    }
  for (int i = 0; i < ntilde; i++)
    if (d[i] / nsize >= 0.001) {
      hi = x[i] * a + b;
      break;
    }
  for (int k = 0; k < ntilde; k++) {
    if (energies) energies[k] = x[k] * a + b;
    if (dos) dos[k] = d[k];
  }
  if (mu) *mu = x[imu] * a + b;
  if (emin) *emin = lo;
  if (emax) *emax = hi;
  return LTB_OK;
}

int ltb_write_dos(const char* path, int n, const double* energies, const double* dos) {
  if (!path || n < 1 || !energies || !dos) return LTB_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return LTB_ERR_ARG;
  for (int i = 0; i < n; i++) std::fprintf(f, "%lf, %lf  \n", energies[i], dos[i]);  // dosfile.txt (:1714-1717)
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? LTB_OK : LTB_ERR_ARG;
}

}  // extern "C"
