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
// reproducible; the estimator is the same).  The stream is random access
// (normal i uses draws 2(i/2)+1 and 2(i/2)+2), so the seeded entry points draw
// their vectors on the device, straight into the recursion's row-major block;
// they equal ltb_random_vectors' host draws up to the last-ulp differences of
// the device and host log / sqrt / sin / cos.  The moments of all vectors are
// one block recursion on the device (the exact-trace kernels with explicit
// start vectors and moment doubling); P r is getKPE on the vector block, and
// the diagonal estimate is formed on the device as well.
#include <hip/hip_runtime.h>

#include <algorithm>
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

// normal number i of the stream of `seed` (ltb::gaussian_fill's sequence):
// uniforms ((z >> 11) + 1) 2^-53 of SplitMix64 draws 2p+1 and 2p+2, p = i/2,
// then r cos t for even i and r sin t for odd i
__device__ __forceinline__ double gauss_at(unsigned long long seed, unsigned long long i) {
  const unsigned long long p = i >> 1;
  double u[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    unsigned long long z = seed + (2 * p + 1 + q) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    u[q] = (double)((z >> 11) + 1) * 0x1.0p-53;
  }
  const double r = sqrt(-2.0 * log(u[0])), t = 2.0 * kPi * u[1];
  return (i & 1) ? r * sin(t) : r * cos(t);
}

// x[j][k] = vector k's orbital j = normal k*norb + j of the stream, in the
// recursion's precision; columns nvec..ld-1 zero
template <typename T>
__global__ __launch_bounds__(256) void k_gauss_block(unsigned long long seed, int norb, int nvec, int ld,
                                                     T* __restrict__ x) {
  const int k = blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (j >= norb || k >= ld) return;
  x[(size_t)j * ld + k] = k < nvec ? (T)gauss_at(seed, (unsigned long long)k * norb + j) : T(0);
}

// |r_k|^2 of the vectors as the recursion holds them (rounded to T)
template <typename T>
__global__ __launch_bounds__(256) void k_gauss_norm2(unsigned long long seed, int norb, double* __restrict__ r2) {
  const int k = blockIdx.x;
  double s = 0;
  for (int j = threadIdx.x; j < norb; j += 256) {
    const double x = (double)(T)gauss_at(seed, (unsigned long long)k * norb + j);
    s += x * x;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) r2[k] = sh[0];
}

// ratio: diagP_j = sum_k r_jk (P r)_jk / sum_k r_jk^2 (computeDiagonalElements
// :1338-1392); otherwise the numerator alone, whose sum over j is
// sum_k r_k^T P r_k (computeAlphas' trace, :1240-1275)
__global__ __launch_bounds__(256) void k_stoch_diag(unsigned long long seed, int norb, int nvec, int ld,
                                                    const double* __restrict__ theta, int ratio,
                                                    double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= norb) return;
  double num = 0, den = 0;
  for (int k = 0; k < nvec; k++) {
    const double r = gauss_at(seed, (unsigned long long)k * norb + j);
    num += r * theta[(size_t)j * ld + k];
    den += r * r;
  }
  out[j] = ratio ? num / den : num;
}

// P r for nvec seeded vectors (getKPE on the device-drawn block), then per
// orbital the ratio estimate of diag(P) or the numerator of the trace
int stochastic_apply(ltb_ctx* c, int M, const double* w, int nvec, unsigned long long seed, int precision, int ratio,
                     std::vector<double>& out) {
  const int norb = c->norb, ld = (nvec + 63) / 64 * 64;
  const bool f64 = precision == LTB_FP64;
  const size_t n = (size_t)norb * ld;
  DevBuf a, b, th, d;
  LTB_HIP(c, hipMalloc(&a.p, n * (f64 ? 8 : 4)));
  LTB_HIP(c, hipMalloc(&b.p, n * (f64 ? 8 : 4)));
  LTB_HIP(c, hipMalloc(&th.p, n * sizeof(double)));
  LTB_HIP(c, hipMalloc(&d.p, sizeof(double) * norb));
  const dim3 grid(ld / 64, (norb + 3) / 4);
  if (f64)
    hipLaunchKernelGGL(k_gauss_block<double>, grid, dim3(256), 0, c->stream, seed, norb, nvec, ld, (double*)a.p);
  else
    hipLaunchKernelGGL(k_gauss_block<float>, grid, dim3(256), 0, c->stream, seed, norb, nvec, ld, (float*)a.p);
  LTB_HIP(c, hipGetLastError());
  int st = kpe_apply_device(c, M, w, ld, a.p, b.p, (double*)th.p, precision);
  if (st) {
    (void)hipStreamSynchronize(c->stream);  // the buffers are freed on return
    return st;
  }
  hipLaunchKernelGGL(k_stoch_diag, dim3((norb + 255) / 256), dim3(256), 0, c->stream, seed, norb, nvec, ld,
                     (const double*)th.p, ratio, (double*)d.p);
  LTB_HIP(c, hipGetLastError());
  out.resize(norb);
  LTB_HIP(c, hipMemcpyAsync(out.data(), d.p, sizeof(double) * norb, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
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
  if (!c || M < 1 || nvec < 1 || (precision != LTB_FP64 && precision != LTB_FP32_ACC64)) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  const int norb = c->norb, ld = kpe_block_width(c, precision, nvec);
  const bool f64 = precision == LTB_FP64;
  int st;
  if ((st = ensure_buffer(c, &c->d_xa, &c->xabytes, (size_t)norb * ld * (f64 ? 8 : 4)))) return st;
  if ((st = ensure_buffer(c, &c->d_xb, &c->xbbytes, (size_t)norb * ld * (f64 ? 8 : 4)))) return st;
  DevBuf d_r2, d_mom;
  LTB_HIP(c, hipMalloc(&d_r2.p, sizeof(double) * ld));
  LTB_HIP(c, hipMalloc(&d_mom.p, sizeof(double) * (size_t)(M + 1) * ld));
  const dim3 grid(ld / 64, (norb + 3) / 4);
  if (f64) {
    hipLaunchKernelGGL(k_gauss_block<double>, grid, dim3(256), 0, c->stream, seed, norb, nvec, ld, (double*)c->d_xa);
    hipLaunchKernelGGL(k_gauss_norm2<double>, dim3(nvec), dim3(256), 0, c->stream, seed, norb, (double*)d_r2.p);
  } else {
    hipLaunchKernelGGL(k_gauss_block<float>, grid, dim3(256), 0, c->stream, seed, norb, nvec, ld, (float*)c->d_xa);
    hipLaunchKernelGGL(k_gauss_norm2<float>, dim3(nvec), dim3(256), 0, c->stream, seed, norb, (double*)d_r2.p);
  }
  LTB_HIP(c, hipGetLastError());
  if ((st = kpe_block_moments_impl(c, M, nvec, precision, (const double*)d_r2.p, (double*)d_mom.p, ld))) return st;
  std::vector<double> mom((size_t)(M + 1) * ld);
  LTB_HIP(c, hipMemcpyAsync(mom.data(), d_mom.p, sizeof(double) * mom.size(), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  if (trace)
    for (int m = 0; m <= M; m++) {
      double s = 0;
      for (int k = 0; k < nvec; k++) s += mom[(size_t)m * ld + k];
      trace[m] = s;
    }
  return LTB_OK;
}

int ltb_density_stochastic(ltb_ctx* c, int M, const double* w, int nvec, unsigned long long seed,
                           int precision, double* diagP) {
  if (!c || M < 1 || !w || !diagP || nvec < 1 || (precision != LTB_FP64 && precision != LTB_FP32_ACC64))
    return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  std::vector<double> d;
  const int st = stochastic_apply(c, M, w, nvec, seed, precision, 1, d);
  if (st) return st;
  std::copy(d.begin(), d.end(), diagP);
  return LTB_OK;
}

int ltb_trace_stochastic(ltb_ctx* c, int M, const double* w, int nvec, unsigned long long seed, int precision,
                         double* trace) {
  if (!c || M < 1 || !w || !trace || nvec < 1 || (precision != LTB_FP64 && precision != LTB_FP32_ACC64))
    return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  std::vector<double> num;
  const int st = stochastic_apply(c, M, w, nvec, seed, precision, 0, num);
  if (st) return st;
  double s = 0;
  for (double x : num) s += x;
  *trace = s / nvec;
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
