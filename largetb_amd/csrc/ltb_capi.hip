// C ABI of the density-matrix path: host structure objects, the device
// context, and the FP64 kernels around the Chebyshev recursion (Gershgorin
// bounds, rescale, Coulomb shift, Fermi-Dirac coefficients, mu bisection,
// populations, band energy, charge mixing).  Reference: pghale/largeTB
// withSP2_deltaH.cpp; each kernel cites the function it replaces.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>
#include <utility>

#include "ctx.hpp"
#include "system.hpp"

namespace {
thread_local std::string g_sys_err;
constexpr double kPi = 3.141592653589793238463;
constexpr int kRedBlocks = 1024;

int sys_fail(const ltb::Status& st) {
  g_sys_err = st.msg;
  return st.code;
}

// ---- block reductions (deterministic fixed trees) ---------------------------------
template <int NT>
__device__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = NT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// ---- constructdeltaH (:1414-1487): v = Gamma (p - p0) ------------------------------
__global__ void k_gamma_shift(const int* __restrict__ grow, const int* __restrict__ gcol,
                              const double* __restrict__ gval, const double* __restrict__ p,
                              const double* __restrict__ p0, double* __restrict__ v, int nat) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nat) return;
  double s = 0;
  for (int k = grow[i]; k < grow[i + 1]; k++) s += gval[k] * (p[gcol[k]] - p0[gcol[k]]);
  v[i] = s;
}

// ---- GershGorinParallel (:144-209) on H0 + diag(v) ---------------------------------
__global__ __launch_bounds__(256) void k_gersh_partial(const double* __restrict__ diag0,
                                                       const double* __restrict__ offabs,
                                                       const double* __restrict__ v, int norb,
                                                       double* __restrict__ red) {
  double lo = INFINITY, hi = -INFINITY;
  int bad = 0;
  for (int r = blockIdx.x * 256 + threadIdx.x; r < norb; r += gridDim.x * 256) {
    double d = diag0[r] + v[r >> 2];
    double o = offabs[r];
    if (!isfinite(d) || !isfinite(o)) bad = 1;
    lo = fmin(lo, d - o);
    hi = fmax(hi, d + o);
  }
  __shared__ double sl[256], shh[256];
  __shared__ int sb;
  if (threadIdx.x == 0) sb = 0;
  __syncthreads();
  if (bad) atomicOr(&sb, 1);
  sl[threadIdx.x] = lo;
  shh[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + w]);
      shh[threadIdx.x] = fmax(shh[threadIdx.x], shh[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[3 * blockIdx.x] = sl[0];
    red[3 * blockIdx.x + 1] = shh[0];
    red[3 * blockIdx.x + 2] = sb ? 1.0 : 0.0;
  }
}

__global__ void k_gersh_final(const double* __restrict__ red, int nb, DevScalars* sc) {
  if (threadIdx.x != 0) return;
  double lo = INFINITY, hi = -INFINITY;
  int bad = 0;
  for (int b = 0; b < nb; b++) {
    lo = fmin(lo, red[3 * b]);
    hi = fmax(hi, red[3 * b + 1]);
    if (red[3 * b + 2] != 0.0) bad = 1;
  }
  if (!isfinite(lo) || !isfinite(hi)) bad = 1;
  sc->emin = lo;
  sc->emax = hi;
  sc->a = (hi - lo) / 2;  // runKPM :1551-1552
  sc->b = (hi + lo) / 2;
  sc->nonfinite |= bad;
}

// ---- runKPM rescale (:1559-1560): H~ = (H0 + diag(v) - b) / a ---------------------
__global__ void k_rescale(const double* __restrict__ h0, const int* __restrict__ blkrow,
                          const int* __restrict__ col, const double* __restrict__ v,
                          const DevScalars* __restrict__ sc, int nnzb, double* __restrict__ ht64,
                          float* __restrict__ ht32) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nnzb * 16) return;
  int k = idx >> 4, e = idx & 15;
  double val = h0[idx];
  int row = blkrow[k];
  if (col[k] == row && (e % 5) == 0) val = (val + v[row]) - sc->b;
  val = val / sc->a;
  ht64[idx] = val;
  ht32[idx] = (float)val;
}

// ---- Fermi-Dirac coefficients and mu bisection (replaces runKPM :1562-1712) -------
__device__ __forceinline__ double fermi(double x, double mu, double kT) {
  double t = (x - mu) / kT;
  if (t > 0) {
    double e = exp(-t);
    return e / (1.0 + e);
  }
  return 1.0 / (1.0 + exp(t));
}

// DOS weights at the K Chebyshev-Gauss nodes: the electron count per spin is
// sum_k f(x_k) W_k = sum_m w_m mu_m with w from the K-node quadrature.
__global__ void k_fd_weights(const double* __restrict__ tot, const double* __restrict__ g, int M,
                             int K, double* __restrict__ W, double* __restrict__ X) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double th = kPi * (k + 0.5) / K;
  double s = tot[0];
  for (int m = 1; m < M; m++) s += 2.0 * g[m] * tot[m] * cos(m * th);
  W[k] = s / K;
  X[k] = cos(th);
}

// the K Chebyshev-Gauss nodes alone (k_fd_weights writes them along with W)
__global__ void k_nodes(int K, double* __restrict__ X) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) X[k] = cos(kPi * (k + 0.5) / K);
}

__device__ double step_count(double mu, const double* g, const double* tot, int M) {
  // getKPE mu_mvals (:1100-1103) times gmMs, summed against the moments
  double am = acos(mu);
  double s = (1 - am / kPi) * tot[0];
  for (int m = 1; m < M; m++) s += (-2 * sin(m * am) / (m * kPi)) * g[m] * tot[m];
  return s;
}

__global__ __launch_bounds__(1024) void k_bisect(const double* __restrict__ W,
                                                 const double* __restrict__ X, int K,
                                                 const double* __restrict__ g,
                                                 const double* __restrict__ tot, int M, double kT,
                                                 double Ne, DevScalars* sc) {
  __shared__ double sh[1024];
  __shared__ double s_lo, s_hi;
  const double kTs = kT / sc->a;
  if (threadIdx.x == 0) {
    s_lo = -1.0;
    s_hi = 1.0;
  }
  __syncthreads();
  for (int it = 0; it < 200; it++) {
    double lo = s_lo, hi = s_hi;
    double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double part = 0;
    if (kT > 0) {
      for (int k = threadIdx.x; k < K; k += 1024) part += fermi(X[k], mid, kTs) * W[k];
    } else if (threadIdx.x == 0) {
      part = step_count(mid, g, tot, M);
    }
    double n = block_sum<1024>(part, sh);
    if (threadIdx.x == 0) {
      if (2.0 * n < Ne)
        s_lo = mid;
      else
        s_hi = mid;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sc->mu = 0.5 * (s_lo + s_hi);
}

// c_m = (2/K) sum_k f(x_k) cos(m th_k); w_0 = c_0/2, w_m = c_m g_m (getKPE :1136-1166)
__global__ __launch_bounds__(256) void k_coeffs(const double* __restrict__ X, int K,
                                                const double* __restrict__ g, int M, double kT,
                                                const DevScalars* __restrict__ sc,
                                                double* __restrict__ w) {
  __shared__ double sh[256];
  const int m = blockIdx.x;
  const double mu = sc->mu;
  double c;
  if (kT > 0) {
    const double kTs = kT / sc->a;
    double s = 0;
    for (int k = threadIdx.x; k < K; k += 256) {
      double th = kPi * (k + 0.5) / K;
      s += fermi(X[k], mu, kTs) * cos(m * th);
    }
    c = block_sum<256>(s, sh) * (2.0 / K);
  } else {
    double am = acos(mu);
    c = m == 0 ? 2 * (1 - am / kPi) : -2 * sin(m * am) / (m * kPi);
  }
  if (threadIdx.x == 0) w[m] = m == 0 ? c / 2.0 : c * g[m];
}

// diag(P)_j = sum_m w_m [T_m]_jj (computeDiagonalElements exact mode, :1361-1379)
__global__ void k_diagp(const double* __restrict__ momorb, const double* __restrict__ w, int M,
                        int norb, double* __restrict__ d) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= norb) return;
  double s = 0;
  for (int m = 0; m < M; m++) s += w[m] * momorb[(size_t)m * norb + j];
  d[j] = s;
}

// computeDiagonalElements (exact) + orbital2Atom (:1385): p_A = sum_m w_m d_m(A)
__global__ void k_populations(const double* __restrict__ mom, const double* __restrict__ w, int M,
                              int nat, double* __restrict__ pop) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat) return;
  double s = 0;
  for (int m = 0; m < M; m++) s += w[m] * mom[(size_t)m * nat + a];
  pop[a] = s;
}

// electrons and band energy 2 Tr(P H) with T_m(x) x = (T_{m+1} + T_{|m-1|}) / 2
__global__ void k_eband(const double* __restrict__ tot, const double* __restrict__ w, int M,
                        DevScalars* sc) {
  if (threadIdx.x != 0) return;
  double ns = 0, tph = w[0] * tot[1];
  for (int m = 0; m < M; m++) ns += w[m] * tot[m];
  for (int m = 1; m < M; m++) tph += w[m] * 0.5 * (tot[m + 1] + tot[m - 1]);
  sc->nelec = 2.0 * ns;
  sc->eband = 2.0 * (sc->a * tph + sc->b * ns);
  if (!isfinite(sc->eband) || !isfinite(sc->mu)) sc->nonfinite = 1;
}

// |q_out - q_in|_inf (e), and the first-order band-energy correction
// 2 sum_A pop_A [Gamma (p_out - p_in)]_A: E_band depends on p_in through the
// Coulomb shift with gradient 2 Gamma pop (Hellmann-Feynman, gap >> kT), so
// E_band(p_in) + 2 pop . Gamma (p_out - p_in) estimates E_band at the fixed
// point to O(|p_out - p*|).
__global__ __launch_bounds__(256) void k_dq_partial(const double* __restrict__ pin,
                                                    const double* __restrict__ pout,
                                                    const int* __restrict__ grow,
                                                    const int* __restrict__ gcol,
                                                    const double* __restrict__ gval, int nat,
                                                    double* __restrict__ red) {
  double mx = 0, corr = 0;
  int bad = 0;
  for (int a = blockIdx.x * 256 + threadIdx.x; a < nat; a += gridDim.x * 256) {
    double d = 2.0 * fabs(pout[a] - pin[a]);
    if (!isfinite(d)) bad = 1;
    mx = fmax(mx, d);
    double gd = 0;
    for (int k = grow[a]; k < grow[a + 1]; k++) gd += gval[k] * (pout[gcol[k]] - pin[gcol[k]]);
    corr += pout[a] * gd;
  }
  __shared__ double sh[256], sc2[256];
  __shared__ int sb;
  if (threadIdx.x == 0) sb = 0;
  __syncthreads();
  if (bad) atomicOr(&sb, 1);
  sh[threadIdx.x] = mx;
  sc2[threadIdx.x] = corr;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
      sc2[threadIdx.x] += sc2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    red[3 * blockIdx.x] = sh[0];
    red[3 * blockIdx.x + 1] = sb ? 1.0 : 0.0;
    red[3 * blockIdx.x + 2] = sc2[0];
  }
}

__global__ void k_dq_final(const double* __restrict__ red, int nb, DevScalars* sc) {
  if (threadIdx.x != 0) return;
  double mx = 0, corr = 0;
  int bad = 0;
  for (int b = 0; b < nb; b++) {
    mx = fmax(mx, red[3 * b]);
    if (red[3 * b + 1] != 0.0) bad = 1;
    corr += red[3 * b + 2];
  }
  sc->dq = mx;
  sc->eband_fp = sc->eband + 2.0 * corr;
  sc->nonfinite |= bad;
}

__global__ void k_mix(double* __restrict__ pin, const double* __restrict__ pout, double mix,
                      int nat) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < nat) pin[a] += mix * (pout[a] - pin[a]);
}

// Gamma_AA of every atom (the diagonal entry of its CSR row)
__global__ void k_gamma_onsite(const int* __restrict__ grow, const int* __restrict__ gcol,
                               double* __restrict__ gval, const double* __restrict__ u, int nat) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat) return;
  for (int k = grow[a]; k < grow[a + 1]; k++)
    if (gcol[k] == a) gval[k] = u[a];
}

__global__ void k_clear_flag(DevScalars* sc) {
  if (threadIdx.x == 0) sc->nonfinite = 0;
}

// two-level SCC: keep the low-order populations, form the defect, apply it
__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int n) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) dst[a] = src[a];
}

__global__ void k_defect(double* __restrict__ d, const double* __restrict__ pout,
                         const double* __restrict__ plo, int n) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) d[a] = pout[a] - plo[a];
}

__global__ void k_add_defect(double* __restrict__ pout, const double* __restrict__ d, int n) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < n) pout[a] += d[a];
}

// gmMs kernel of getKPE (:1105-1132)
void kernel_gmM(int M, std::vector<double>& g) {
  std::vector<double> U(M);
  double alpha = kPi / (M + 2), lam = std::cos(alpha), nrm = 0;
  for (int m = 0; m < M; m++) {
    U[m] = std::sin((m + 1) * std::acos(lam)) / std::sin(std::acos(lam));
    nrm += U[m] * U[m];
  }
  nrm = std::sqrt(nrm);
  for (int i = 0; i < M; i++) U[i] /= nrm;
  g.assign(M, 0.0);
  for (int m = 0; m < M; m++) {
    double s = 0;
    for (int i = 0; i < M - m - 1; i++) s += U[i] * U[m + i];
    g[m] = s;
  }
}

int need_system(ltb_ctx* c) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  return LTB_OK;
}

int grid_for(int n, int bs) { return (n + bs - 1) / bs; }

// Fermi quadrature / kernel setup for (M, K)
int ensure_fermi(ltb_ctx* c, int M, int K) {
  if (c->fd_M == M && c->fd_K == K && c->d_W && c->d_X && c->d_g && c->d_w) return LTB_OK;
  (void)hipFree(c->d_W);
  (void)hipFree(c->d_X);
  (void)hipFree(c->d_g);
  (void)hipFree(c->d_w);
  c->d_W = c->d_X = c->d_g = c->d_w = nullptr;
  c->fd_M = c->fd_K = 0;  // a failed allocation below leaves no set that looks valid
  LTB_HIP(c, hipMalloc(&c->d_W, sizeof(double) * std::max(K, 1)));
  LTB_HIP(c, hipMalloc(&c->d_X, sizeof(double) * std::max(K, 1)));
  LTB_HIP(c, hipMalloc(&c->d_g, sizeof(double) * M));
  LTB_HIP(c, hipMalloc(&c->d_w, sizeof(double) * M));
  std::vector<double> g;
  kernel_gmM(M, g);
  LTB_HIP(c, hipMemcpy(c->d_g, g.data(), sizeof(double) * M, hipMemcpyHostToDevice));
  c->fd_M = M;
  c->fd_K = K;
  return LTB_OK;
}

// density step from the moments already on the device (no host sync)
int density_async(ltb_ctx* c, const ltb_fermi_params* fp) {
  const int M = fp->M, K = fp->kT > 0 ? fp->K : 1;
  if (c->mom_M < M) return set_err(c, LTB_ERR_STATE, "moments missing: run ltb_kpe_moments with M first");
  int st = ensure_fermi(c, M, K);
  if (st) return st;
  if (fp->kT > 0)
    hipLaunchKernelGGL(k_fd_weights, dim3(grid_for(K, 256)), dim3(256), 0, c->stream, c->d_momtot,
                       c->d_g, M, K, c->d_W, c->d_X);
  hipLaunchKernelGGL(k_bisect, dim3(1), dim3(1024), 0, c->stream, c->d_W, c->d_X, K, c->d_g,
                     c->d_momtot, M, fp->kT, fp->Ne, c->d_sc);
  hipLaunchKernelGGL(k_coeffs, dim3(M), dim3(256), 0, c->stream, c->d_X, K, c->d_g, M, fp->kT,
                     c->d_sc, c->d_w);
  hipLaunchKernelGGL(k_populations, dim3(grid_for(c->nat, 256)), dim3(256), 0, c->stream, c->d_mom,
                     c->d_w, M, c->nat, c->d_pout);
  hipLaunchKernelGGL(k_eband, dim3(1), dim3(64), 0, c->stream, c->d_momtot, c->d_w, M, c->d_sc);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int shift_async(ltb_ctx* c) {
  hipLaunchKernelGGL(k_gamma_shift, dim3(grid_for(c->nat, 256)), dim3(256), 0, c->stream, c->d_grow,
                     c->d_gcol, c->d_gval, c->d_pin, c->d_p0, c->d_v, c->nat);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int gersh_async(ltb_ctx* c) {
  int nb = std::min(kRedBlocks, grid_for(c->norb, 256));
  hipLaunchKernelGGL(k_gersh_partial, dim3(nb), dim3(256), 0, c->stream, c->d_diag0, c->d_offabs,
                     c->d_v, c->norb, c->d_red);
  hipLaunchKernelGGL(k_gersh_final, dim3(1), dim3(64), 0, c->stream, c->d_red, nb, c->d_sc);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int rescale_async(ltb_ctx* c) {
  hipLaunchKernelGGL(k_rescale, dim3(grid_for(c->nnzb * 16, 256)), dim3(256), 0, c->stream, c->d_h0,
                     c->d_blkrow, c->d_col, c->d_v, c->d_sc, c->nnzb, c->d_ht64, c->d_ht32);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int mix_async(ltb_ctx* c, double mix) {
  int nb = std::min(kRedBlocks, grid_for(c->nat, 256));
  hipLaunchKernelGGL(k_dq_partial, dim3(nb), dim3(256), 0, c->stream, c->d_pin, c->d_pout,
                     c->d_grow, c->d_gcol, c->d_gval, c->nat, c->d_red);
  hipLaunchKernelGGL(k_dq_final, dim3(1), dim3(64), 0, c->stream, c->d_red, nb, c->d_sc);
  hipLaunchKernelGGL(k_mix, dim3(grid_for(c->nat, 256)), dim3(256), 0, c->stream, c->d_pin,
                     c->d_pout, mix, c->nat);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int read_scalars(ltb_ctx* c, DevScalars* h) {
  LTB_HIP(c, hipMemcpyAsync(h, c->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

void free_all(ltb_ctx* c) {
  void* ptrs[] = {c->d_rowptr, c->d_col, c->d_blkrow, c->d_h0, c->d_ht64, c->d_ht32, c->d_offabs,
                  c->d_diag0, c->d_grow, c->d_gcol, c->d_gval, c->d_p0, c->d_pin, c->d_pout,
                  c->d_v, c->d_sc, c->d_red, c->d_xa, c->d_xb, c->d_part, c->d_mu1, c->d_mom,
                  c->d_W, c->d_X, c->d_g, c->d_c, c->d_w, c->d_sa, c->d_sb, c->d_momorb, c->d_visit,
                  c->lo_W, c->lo_X, c->lo_g, c->lo_w, c->d_plo, c->d_defect};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  free_cones(c);
}

}  // namespace

int set_err(ltb_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int ensure_buffer(ltb_ctx* c, void** p, size_t* cap, size_t bytes) {
  if (*p && *cap >= bytes) return LTB_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, LTB_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  *cap = bytes;
  return LTB_OK;
}

extern "C" {

// ---------------------------- host structure -------------------------------------
int ltb_system_quartz(const ltb_quartz_desc* d, ltb_system** out) {
  if (!d || !out) return sys_fail({LTB_ERR_ARG, "null argument"});
  auto* s = new ltb_system();
  ltb::Status st = ltb::build_quartz(d->L, d->M, d->N, d->lattice, d->sigma_ang, d->seed, *s);
  if (st.code) {
    delete s;
    *out = nullptr;
    return sys_fail(st);
  }
  *out = s;
  return LTB_OK;
}

int ltb_system_create(int natoms, const int* Z, const double* xyz, const double* cell9, int pmask,
                      ltb_system** out) {
  if (!out || natoms <= 0 || !Z || !xyz) return sys_fail({LTB_ERR_ARG, "bad arguments"});
  auto* s = new ltb_system();
  s->natoms = natoms;
  s->Z.assign(Z, Z + natoms);
  s->xyz.assign(xyz, xyz + 3 * (size_t)natoms);
  if (cell9)
    std::memcpy(s->cell, cell9, sizeof(double) * 9);
  else if (pmask)
    return delete s, sys_fail({LTB_ERR_ARG, "periodic_mask needs a cell"});
  s->periodic_mask = pmask & 7;
  ltb::Status st = ltb::build_from_coords(*s);
  if (st.code) {
    delete s;
    *out = nullptr;
    return sys_fail(st);
  }
  *out = s;
  return LTB_OK;
}

void ltb_system_destroy(ltb_system* s) { delete s; }

const char* ltb_system_last_error(void) { return g_sys_err.c_str(); }

int ltb_system_sizes(const ltb_system* s, int* nat, int* nnzb, int* gnnz, double* ne) {
  if (!s) return sys_fail({LTB_ERR_ARG, "null system"});
  if (nat) *nat = s->natoms;
  if (nnzb) *nnzb = (int)s->h_col.size();
  if (gnnz) *gnnz = (int)s->g_col.size();
  if (ne) *ne = s->nelectrons;
  return LTB_OK;
}

int ltb_system_coords(const ltb_system* s, int* Z, double* xyz) {
  if (!s) return sys_fail({LTB_ERR_ARG, "null system"});
  if (Z) std::memcpy(Z, s->Z.data(), sizeof(int) * s->natoms);
  if (xyz) std::memcpy(xyz, s->xyz.data(), sizeof(double) * 3 * s->natoms);
  return LTB_OK;
}

int ltb_system_hamiltonian(const ltb_system* s, int* rowptr, int* col, double* val) {
  if (!s) return sys_fail({LTB_ERR_ARG, "null system"});
  if (rowptr) std::memcpy(rowptr, s->h_rowptr.data(), sizeof(int) * s->h_rowptr.size());
  if (col) std::memcpy(col, s->h_col.data(), sizeof(int) * s->h_col.size());
  if (val) std::memcpy(val, s->h_val.data(), sizeof(double) * s->h_val.size());
  return LTB_OK;
}

int ltb_system_gamma(const ltb_system* s, int* rowptr, int* col, double* val) {
  if (!s) return sys_fail({LTB_ERR_ARG, "null system"});
  if (rowptr) std::memcpy(rowptr, s->g_rowptr.data(), sizeof(int) * s->g_rowptr.size());
  if (col) std::memcpy(col, s->g_col.data(), sizeof(int) * s->g_col.size());
  if (val) std::memcpy(val, s->g_val.data(), sizeof(double) * s->g_val.size());
  return LTB_OK;
}

int ltb_system_neutral(const ltb_system* s, double* p0) {
  if (!s || !p0) return sys_fail({LTB_ERR_ARG, "null argument"});
  std::memcpy(p0, s->p0.data(), sizeof(double) * s->natoms);
  return LTB_OK;
}

// ---------------------------- device context -------------------------------------
int ltb_create(int device, ltb_ctx** out) {
  if (!out) return LTB_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    g_sys_err = "no HIP device";
    return LTB_ERR_DEVICE;
  }
  if (device < 0 || device >= n) {
    g_sys_err = "device index out of range";
    return LTB_ERR_ARG;
  }
  auto* c = new ltb_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    g_sys_err = "stream creation failed";
    return LTB_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
  if (hipMalloc(&c->d_sc, sizeof(DevScalars)) != hipSuccess ||
      hipMalloc(&c->d_red, sizeof(double) * 3 * kRedBlocks) != hipSuccess) {
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    g_sys_err = "allocation failed";
    return LTB_ERR_NOMEM;
  }
  if (hipMemset(c->d_sc, 0, sizeof(DevScalars)) != hipSuccess) {
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    g_sys_err = "device scalar initialisation failed";
    return LTB_ERR_DEVICE;
  }
  if (noalloc_scratch_init(c) != LTB_OK) {
    g_sys_err = c->err;
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return LTB_ERR_DEVICE;
  }
  *out = c;
  return LTB_OK;
}

void ltb_destroy(ltb_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_all(c);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* ltb_last_error(const ltb_ctx* c) { return c ? c->err.c_str() : g_sys_err.c_str(); }

}  // extern "C"

// Drop the buffers sized by the loaded system.
void release_system(ltb_ctx* c) {
  void** ps[] = {(void**)&c->d_rowptr, (void**)&c->d_col, (void**)&c->d_blkrow, (void**)&c->d_h0,
                 (void**)&c->d_ht64, (void**)&c->d_ht32, (void**)&c->d_offabs, (void**)&c->d_diag0,
                 (void**)&c->d_grow, (void**)&c->d_gcol, (void**)&c->d_gval, (void**)&c->d_p0,
                 (void**)&c->d_pin, (void**)&c->d_pout, (void**)&c->d_v, (void**)&c->d_visit,
                 (void**)&c->d_plo, (void**)&c->d_defect};
  for (void** p : ps) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  c->mom_M = -1;
  c->nat = c->norb = c->nnzb = c->gnnz = 0;
  c->nelectrons = 0;
  free_cones(c);
  c->h_rowptr.clear();
  c->h_col.clear();
}

// Allocate the buffers of a system with nat atoms, nnzb H blocks and gnnz
// Gamma entries (contents not set).
int alloc_system(ltb_ctx* c, int nat, int nnzb, int gnnz) {
  const int norb = 4 * nat;
  LTB_HIP(c, hipMalloc(&c->d_rowptr, sizeof(int) * (nat + 1)));
  LTB_HIP(c, hipMalloc(&c->d_col, sizeof(int) * nnzb));
  LTB_HIP(c, hipMalloc(&c->d_blkrow, sizeof(int) * nnzb));
  LTB_HIP(c, hipMalloc(&c->d_h0, sizeof(double) * 16 * nnzb));
  LTB_HIP(c, hipMalloc(&c->d_ht64, sizeof(double) * 16 * nnzb));
  LTB_HIP(c, hipMalloc(&c->d_ht32, sizeof(float) * 16 * nnzb));
  LTB_HIP(c, hipMalloc(&c->d_offabs, sizeof(double) * norb));
  LTB_HIP(c, hipMalloc(&c->d_diag0, sizeof(double) * norb));
  LTB_HIP(c, hipMalloc(&c->d_grow, sizeof(int) * (nat + 1)));
  LTB_HIP(c, hipMalloc(&c->d_gcol, sizeof(int) * gnnz));
  LTB_HIP(c, hipMalloc(&c->d_gval, sizeof(double) * gnnz));
  LTB_HIP(c, hipMalloc(&c->d_p0, sizeof(double) * nat));
  LTB_HIP(c, hipMalloc(&c->d_pin, sizeof(double) * nat));
  LTB_HIP(c, hipMalloc(&c->d_pout, sizeof(double) * nat));
  LTB_HIP(c, hipMalloc(&c->d_v, sizeof(double) * nat));
  LTB_HIP(c, hipMalloc(&c->d_visit, sizeof(int) * nat));
  LTB_HIP(c, hipMalloc(&c->d_plo, sizeof(double) * nat));
  LTB_HIP(c, hipMalloc(&c->d_defect, sizeof(double) * nat));
  LTB_HIP(c, hipMemset(c->d_v, 0, sizeof(double) * nat));
  return LTB_OK;
}

extern "C" {

int ltb_load_system(ltb_ctx* c, const ltb_system* s) {
  if (!c || !s) return LTB_ERR_ARG;
  LTB_HIP(c, hipSetDevice(c->device));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  release_system(c);
  const int nat = s->natoms, nnzb = (int)s->h_col.size(), gnnz = (int)s->g_col.size();
  const int norb = 4 * nat;
  std::vector<int> blkrow(nnzb);
  std::vector<double> diag0(norb, 0.0), offabs(norb, 0.0);
  for (int i = 0; i < nat; i++)
    for (int k = s->h_rowptr[i]; k < s->h_rowptr[i + 1]; k++) {
      blkrow[k] = i;
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          double v = s->h_val[16 * (size_t)k + 4 * a + b];
          if (s->h_col[k] == i && a == b)
            diag0[4 * i + a] = v;
          else
            offabs[4 * i + a] += std::fabs(v);
        }
    }
  std::vector<int> visit;
  ltb::Status vs = ltb::locality_order(*s, visit);
  if (vs.code) return set_err(c, LTB_ERR_ARG, vs.msg);
  int st = alloc_system(c, nat, nnzb, gnnz);
  if (st) {
    release_system(c);
    return st;
  }
  LTB_HIP(c, hipMemcpy(c->d_rowptr, s->h_rowptr.data(), sizeof(int) * (nat + 1), hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_col, s->h_col.data(), sizeof(int) * nnzb, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_blkrow, blkrow.data(), sizeof(int) * nnzb, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_h0, s->h_val.data(), sizeof(double) * 16 * nnzb, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_offabs, offabs.data(), sizeof(double) * norb, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_diag0, diag0.data(), sizeof(double) * norb, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_grow, s->g_rowptr.data(), sizeof(int) * (nat + 1), hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_gcol, s->g_col.data(), sizeof(int) * gnnz, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_gval, s->g_val.data(), sizeof(double) * gnnz, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_p0, s->p0.data(), sizeof(double) * nat, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_pin, s->p0.data(), sizeof(double) * nat, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(c->d_visit, visit.data(), sizeof(int) * nat, hipMemcpyHostToDevice));
  c->nat = nat;
  c->norb = norb;
  c->nnzb = nnzb;
  c->gnnz = gnnz;
  c->nelectrons = s->nelectrons;
  return LTB_OK;
}

int ltb_set_populations(ltb_ctx* c, const double* p) {
  int st = need_system(c);
  if (st) return st;
  if (p) LTB_HIP(c, hipMemcpyAsync(c->d_pin, p, sizeof(double) * c->nat, hipMemcpyHostToDevice, c->stream));
  if ((st = shift_async(c))) return st;
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_get_shift(ltb_ctx* c, double* v) {
  int st = need_system(c);
  if (st) return st;
  if (!v) return set_err(c, LTB_ERR_ARG, "null output");
  LTB_HIP(c, hipMemcpyAsync(v, c->d_v, sizeof(double) * c->nat, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_set_gamma_onsite(ltb_ctx* c, const double* gamma_aa) {
  int st = need_system(c);
  if (st) return st;
  if (!gamma_aa) return set_err(c, LTB_ERR_ARG, "null onsite Gamma");
  for (int i = 0; i < c->nat; i++)
    if (!std::isfinite(gamma_aa[i])) return set_err(c, LTB_ERR_NONFINITE, "non-finite onsite Gamma");
  double* d = nullptr;
  LTB_HIP(c, hipMalloc(&d, sizeof(double) * c->nat));
  hipError_t e = hipMemcpy(d, gamma_aa, sizeof(double) * c->nat, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_gamma_onsite, dim3(grid_for(c->nat, 256)), dim3(256), 0, c->stream, c->d_grow, c->d_gcol,
                       c->d_gval, d, c->nat);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  (void)hipFree(d);
  if (e != hipSuccess) return set_err(c, LTB_ERR_DEVICE, std::string("onsite Gamma: ") + hipGetErrorString(e));
  return LTB_OK;
}

int ltb_set_neutral(ltb_ctx* c, const double* p0) {
  int st = need_system(c);
  if (st) return st;
  if (!p0) return set_err(c, LTB_ERR_ARG, "null neutral populations");
  for (int i = 0; i < c->nat; i++)
    if (!std::isfinite(p0[i])) return set_err(c, LTB_ERR_NONFINITE, "non-finite neutral population");
  LTB_HIP(c, hipMemcpyAsync(c->d_p0, p0, sizeof(double) * c->nat, hipMemcpyHostToDevice, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_get_mixed_populations(ltb_ctx* c, double* p) {
  int st = need_system(c);
  if (st) return st;
  if (!p) return set_err(c, LTB_ERR_ARG, "null output");
  LTB_HIP(c, hipMemcpyAsync(p, c->d_pin, sizeof(double) * c->nat, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_gershgorin(ltb_ctx* c, double* emin, double* emax) {
  int st = need_system(c);
  if (st) return st;
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  if ((st = gersh_async(c))) return st;
  DevScalars h;
  if ((st = read_scalars(c, &h))) return st;
  if (emin) *emin = h.emin;
  if (emax) *emax = h.emax;
  if (h.nonfinite) return set_err(c, LTB_ERR_NONFINITE, "non-finite entry in H_total");
  return LTB_OK;
}

int ltb_rescale(ltb_ctx* c, double a, double b) {
  int st = need_system(c);
  if (st) return st;
  if (!(a > 0) || !std::isfinite(a) || !std::isfinite(b)) return set_err(c, LTB_ERR_ARG, "rescale needs finite a > 0");
  DevScalars h;
  if ((st = read_scalars(c, &h))) return st;
  h.a = a;
  h.b = b;
  LTB_HIP(c, hipMemcpyAsync(c->d_sc, &h, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  if ((st = rescale_async(c))) return st;
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_kpe_moments(ltb_ctx* c, int M, int precision, double* atom_moments, double* total) {
  int st = need_system(c);
  if (st) return st;
  if (M < 2 || M > 1 << 16) return set_err(c, LTB_ERR_ARG, "M must be in [2, 65536]");
  if (precision != LTB_FP64 && precision != LTB_FP32_ACC64) return set_err(c, LTB_ERR_ARG, "bad precision");
  if ((st = kpe_moments_impl(c, M, precision))) return st;
  c->momorb_M = c->keep_orbital ? M : -1;
  if (atom_moments)
    LTB_HIP(c, hipMemcpyAsync(atom_moments, c->d_mom, sizeof(double) * (size_t)(M + 1) * c->nat,
                              hipMemcpyDeviceToHost, c->stream));
  if (total)
    LTB_HIP(c, hipMemcpyAsync(total, c->d_momtot, sizeof(double) * (M + 1), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_density(ltb_ctx* c, const ltb_fermi_params* fp, double* pop, ltb_density_result* res) {
  int st = need_system(c);
  if (st) return st;
  if (!fp || fp->M < 2 || (fp->kT > 0 && fp->K < 16)) return set_err(c, LTB_ERR_ARG, "bad Fermi parameters");
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  if ((st = density_async(c, fp))) return st;
  if (pop) LTB_HIP(c, hipMemcpyAsync(pop, c->d_pout, sizeof(double) * c->nat, hipMemcpyDeviceToHost, c->stream));
  DevScalars h;
  if ((st = read_scalars(c, &h))) return st;
  if (res) {
    res->mu = h.mu;
    res->mu_ha = h.a * h.mu + h.b;
    res->a = h.a;
    res->b = h.b;
    res->nelec = h.nelec;
    res->eband = h.eband;
  }
  if (h.nonfinite) return set_err(c, LTB_ERR_NONFINITE, "non-finite density result");
  return LTB_OK;
}

// one SCC iteration, device only: constructdeltaH, Gershgorin, rescale,
// exact-trace moments, mu bisection + populations + band energy, dq + mixing
int scc_iteration_async(ltb_ctx* c, const ltb_scc_params* prm, const ltb_fermi_params* fp, int prec) {
  int s2;
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  if ((s2 = shift_async(c)) || (s2 = gersh_async(c)) || (s2 = rescale_async(c))) return s2;
  if ((s2 = kpe_moments_impl(c, prm->M, prec))) return s2;
  if ((s2 = density_async(c, fp))) return s2;
  return mix_async(c, prm->mix);
}

// The low-order density of the two-level SCC runs with the second set of
// quadrature and kernel buffers swapped in (restored on every exit path).
struct LowOrderFermi {
  ltb_ctx* c;
  explicit LowOrderFermi(ltb_ctx* cc) : c(cc) { swap(); }
  ~LowOrderFermi() { swap(); }
  void swap() {
    std::swap(c->d_W, c->lo_W);
    std::swap(c->d_X, c->lo_X);
    std::swap(c->d_g, c->lo_g);
    std::swap(c->d_w, c->lo_w);
    std::swap(c->fd_M, c->lo_M);
    std::swap(c->fd_K, c->lo_K);
  }
};

// the low-order recursion's chunk width (tiling.lo_chunk_cols) while it runs
struct LowOrderChunk {
  ltb_kpe_tiling& t;
  const int saved;
  LowOrderChunk(ltb_ctx* c, int prec) : t(c->tiling[prec == LTB_FP64 ? 1 : 0]), saved(t.chunk_cols) {
    if (t.lo_chunk_cols > 0) t.chunk_cols = t.lo_chunk_cols;
  }
  ~LowOrderChunk() { t.chunk_cols = saved; }
  LowOrderChunk(const LowOrderChunk&) = delete;
  LowOrderChunk& operator=(const LowOrderChunk&) = delete;
};

// populations at order fl->M of the H~ already rescaled at q_in, into d_pout
int low_order_density_async(ltb_ctx* c, const ltb_fermi_params* fl, int prec) {
  LowOrderFermi sw(c);
  LowOrderChunk ch(c, prec);
  int s2;
  if ((s2 = kpe_moments_impl(c, fl->M, prec))) return s2;
  return density_async(c, fl);
}

// two-level SCC, full-order iteration: both orders at q_in, the defect
// between them kept, and the mixing step taken with the full-order output
int scc_full_iteration_async(ltb_ctx* c, const ltb_scc_params* prm, const ltb_fermi_params* fp,
                             const ltb_fermi_params* fl, int prec, int prec_lo) {
  int s2;
  const dim3 g(grid_for(c->nat, 256));
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  if ((s2 = shift_async(c)) || (s2 = gersh_async(c)) || (s2 = rescale_async(c))) return s2;
  if ((s2 = low_order_density_async(c, fl, prec_lo))) return s2;
  hipLaunchKernelGGL(k_copy, g, dim3(256), 0, c->stream, c->d_plo, c->d_pout, c->nat);
  // the full order last, so mu and the band energy on the device are its own
  if ((s2 = kpe_moments_impl(c, prm->M, prec))) return s2;
  if ((s2 = density_async(c, fp))) return s2;
  hipLaunchKernelGGL(k_defect, g, dim3(256), 0, c->stream, c->d_defect, c->d_pout, c->d_plo, c->nat);
  LTB_HIP(c, hipGetLastError());
  return mix_async(c, prm->mix);
}

// two-level SCC, low-order iteration: F_M_lo(q_in) + defect, then mixing
int scc_low_iteration_async(ltb_ctx* c, const ltb_scc_params* prm, const ltb_fermi_params* fl, int prec_lo) {
  int s2;
  hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, c->stream, c->d_sc);
  if ((s2 = shift_async(c)) || (s2 = gersh_async(c)) || (s2 = rescale_async(c))) return s2;
  if ((s2 = low_order_density_async(c, fl, prec_lo))) return s2;
  hipLaunchKernelGGL(k_add_defect, dim3(grid_for(c->nat, 256)), dim3(256), 0, c->stream, c->d_pout,
                     c->d_defect, c->nat);
  LTB_HIP(c, hipGetLastError());
  return mix_async(c, prm->mix);
}

// Two-level SCC (defect correction in the expansion order).  Linear mixing
// contracts the charge error by only (1 - mix) per iteration whatever the
// accuracy of the density, so the density of most iterations can come from
// order M_lo corrected by the defect d = F_M(q_ref) - F_M_lo(q_ref) of the
// last full-order iteration at q_ref: the corrected map F_M_lo(q) + d differs
// from F_M(q) only by (J_M - J_M_lo)(q - q_ref), second order in small
// quantities.  Convergence is declared on full-order iterations only.
int scc_two_level(ltb_ctx* c, const ltb_scc_params* prm, const double* guess, double* p_out, double* hist,
                  ltb_scc_result* res) {
  auto t0 = std::chrono::steady_clock::now();
  int st;
  if (guess)
    LTB_HIP(c, hipMemcpyAsync(c->d_pin, guess, sizeof(double) * c->nat, hipMemcpyHostToDevice, c->stream));
  else
    LTB_HIP(c, hipMemcpyAsync(c->d_pin, c->d_p0, sizeof(double) * c->nat, hipMemcpyDeviceToDevice, c->stream));
  // buffers at their final size before anything is captured
  if ((st = kpe_reserve(c, prm->M, prm->precision))) return st;
  if (prm->switch_tol > 0 && (st = kpe_reserve(c, prm->M, LTB_FP64))) return st;
  {
    LowOrderChunk ch(c, prm->precision);
    if ((st = kpe_reserve(c, prm->M_lo, prm->precision))) return st;
  }
  const ltb_fermi_params fp{prm->M, prm->kT, prm->K, c->nelectrons};
  const ltb_fermi_params fl{prm->M_lo, prm->kT, prm->K, c->nelectrons};
  const double lo_tol = prm->lo_tol > 0 ? prm->lo_tol : 0.1 * prm->tol;
  const int prec_lo = prm->precision;
  int prec = prm->precision;  // of the full-order iterations
  bool full = true;
  int it = 0, n64 = 0, nlo = 0, ngraph = 0, phase_len = 0, stall = 0;
  double best = 0, dq_full = 0;
  DevScalars h{};
  const bool saved_timing = c->timing;
  GraphExecs<1> graph;  // the low-order iteration
  bool warmed = false;
  st = LTB_NOT_CONVERGED;
  for (it = 0; it < prm->max_iter; it++) {
    int s2 = LTB_OK;
    if (full) {
      if ((s2 = scc_full_iteration_async(c, prm, &fp, &fl, prec, prec_lo))) return s2;
      if (prec == LTB_FP64) n64++;
    } else {
      if (prm->use_graph && warmed) {
        if (!graph.e[0]) {
          c->timing = false;
          hipGraph_t g = nullptr;
          LTB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
          s2 = scc_low_iteration_async(c, prm, &fl, prec_lo);
          hipError_t e = hipStreamEndCapture(c->stream, &g);
          if (s2) return s2;
          if (e != hipSuccess) return set_err(c, LTB_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
          e = hipGraphInstantiate(&graph.e[0], g, nullptr, nullptr, 0);
          (void)hipGraphDestroy(g);
          if (e != hipSuccess)
            return set_err(c, LTB_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
        }
        LTB_HIP(c, hipGraphLaunch(graph.e[0], c->stream));
        ngraph++;
      } else {
        if ((s2 = scc_low_iteration_async(c, prm, &fl, prec_lo))) return s2;
        warmed = true;
      }
      nlo++;
      phase_len++;
    }
    if ((s2 = read_scalars(c, &h))) return s2;
    if (h.nonfinite) {
      st = set_err(c, LTB_ERR_NONFINITE, "non-finite value in SCC iteration " + std::to_string(it));
      it++;
      break;
    }
    if (hist) hist[it] = h.dq;
    if (prm->verbose) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "scc iter %d %s M=%d dq=%.3e mu=%.6f Ha eband_fp=%.10f Ha t=%.2fs\n", it,
              (full ? prec : prec_lo) == LTB_FP64 ? "fp64" : "fp32", full ? prm->M : prm->M_lo, h.dq,
              h.a * h.mu + h.b, h.eband_fp, el);
      fflush(stderr);
    }
    if (full) {
      const bool final_precision = prec == LTB_FP64 || prm->switch_tol <= 0;
      if (h.dq < prm->tol && final_precision) {
        st = LTB_OK;
        it++;
        break;
      }
      if (prec != LTB_FP64 && prm->switch_tol > 0 && h.dq < prm->switch_tol) prec = LTB_FP64;
      full = false;
      phase_len = 0;
      dq_full = h.dq;
    } else {
      if (phase_len == 1 || h.dq < 0.9 * best) {
        best = h.dq;
        stall = 0;
      } else {
        stall++;
      }
      if (h.dq < lo_tol || h.dq < kTwoLevelPhaseRel * dq_full || stall >= kTwoLevelStall ||
          phase_len >= kTwoLevelMaxPhase)
        full = true;
    }
  }
  c->timing = saved_timing;
  // d_pout holds the last iteration's output: F_M(q_in) after a full-order one
  if (p_out) LTB_HIP(c, hipMemcpy(p_out, c->d_pout, sizeof(double) * c->nat, hipMemcpyDeviceToHost));
  auto t1 = std::chrono::steady_clock::now();
  if (res) {
    res->iterations = it;
    res->fp64_iterations = n64;
    res->lo_iterations = nlo;
    res->dq = h.dq;
    res->mu_ha = h.a * h.mu + h.b;
    res->eband = h.eband;
    res->eband_fp = h.eband_fp;
    res->graph_launches = ngraph;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
  }
  if (st == LTB_NOT_CONVERGED)
    set_err(c, LTB_NOT_CONVERGED, "SCC not converged after " + std::to_string(prm->max_iter) + " iterations");
  return st;
}

int ltb_scc_run(ltb_ctx* c, const ltb_scc_params* prm, const double* guess, double* p_out,
                double* hist, ltb_scc_result* res) {
  int st = need_system(c);
  if (st) return st;
  if (!prm || prm->M < 2 || prm->max_iter < 1 || !(prm->tol > 0) || !(prm->mix > 0) || prm->mix > 1 ||
      (prm->kT > 0 && prm->K < 16) ||
      (prm->precision != LTB_FP64 && prm->precision != LTB_FP32_ACC64) ||
      prm->M_lo < 0 || (prm->M_lo > 0 && (prm->M_lo < 2 || prm->M_lo >= prm->M)) || !(prm->lo_tol >= 0))
    return set_err(c, LTB_ERR_ARG, "bad SCC parameters");
  if (prm->M_lo > 0) return scc_two_level(c, prm, guess, p_out, hist, res);
  auto t0 = std::chrono::steady_clock::now();
  if (guess)
    LTB_HIP(c, hipMemcpyAsync(c->d_pin, guess, sizeof(double) * c->nat, hipMemcpyHostToDevice, c->stream));
  else
    LTB_HIP(c, hipMemcpyAsync(c->d_pin, c->d_p0, sizeof(double) * c->nat, hipMemcpyDeviceToDevice, c->stream));
  ltb_fermi_params fp{prm->M, prm->kT, prm->K, c->nelectrons};
  int prec = prm->precision;
  int it = 0, n64 = 0, ngraph = 0;
  DevScalars h{};
  const bool saved_timing = c->timing;
  GraphExecs<2> graphs;  // [FP64, FP32]
  hipGraphExec_t* exec = graphs.e;
  bool warmed[2] = {false, false};
  st = LTB_NOT_CONVERGED;
  for (it = 0; it < prm->max_iter; it++) {
    const int gi = prec == LTB_FP64 ? 0 : 1;
    int s2 = LTB_OK;
    if (prm->use_graph && warmed[gi]) {
      if (!exec[gi]) {
        // capture once; buffers were sized by the eager warm iteration
        c->timing = false;
        hipGraph_t g = nullptr;
        LTB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        s2 = scc_iteration_async(c, prm, &fp, prec);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (s2) return s2;
        if (e != hipSuccess) return set_err(c, LTB_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&exec[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return set_err(c, LTB_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
      }
      LTB_HIP(c, hipGraphLaunch(exec[gi], c->stream));
      ngraph++;
    } else {
      c->timing = saved_timing && it == 0;  // per-launch events in the first iteration only
      if ((s2 = scc_iteration_async(c, prm, &fp, prec))) return s2;
      warmed[gi] = true;
    }
    if (prec == LTB_FP64) n64++;
    if ((s2 = read_scalars(c, &h))) return s2;
    if (h.nonfinite) {
      st = set_err(c, LTB_ERR_NONFINITE, "non-finite value in SCC iteration " + std::to_string(it));
      it++;
      break;
    }
    if (hist) hist[it] = h.dq;
    if (prm->verbose) {
      double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "scc iter %d %s dq=%.3e mu=%.6f Ha eband_fp=%.10f Ha t=%.2fs\n", it,
              prec == LTB_FP64 ? "fp64" : "fp32", h.dq, h.a * h.mu + h.b, h.eband_fp, el);
      fflush(stderr);
    }
    const bool final_precision = prec == LTB_FP64 || prm->switch_tol <= 0;
    if (h.dq < prm->tol && final_precision) {
      st = LTB_OK;
      it++;
      break;
    }
    if (prec != LTB_FP64 && prm->switch_tol > 0 && h.dq < prm->switch_tol) prec = LTB_FP64;
  }
  c->timing = saved_timing;
  if (p_out) LTB_HIP(c, hipMemcpy(p_out, c->d_pout, sizeof(double) * c->nat, hipMemcpyDeviceToHost));
  auto t1 = std::chrono::steady_clock::now();
  if (res) {
    res->iterations = it;
    res->fp64_iterations = n64;
    res->dq = h.dq;
    res->mu_ha = h.a * h.mu + h.b;
    res->eband = h.eband;
    res->eband_fp = h.eband_fp;
    res->graph_launches = ngraph;
    res->seconds = std::chrono::duration<double>(t1 - t0).count();
    res->lo_iterations = 0;
  }
  if (st == LTB_NOT_CONVERGED)
    set_err(c, LTB_NOT_CONVERGED, "SCC not converged after " + std::to_string(prm->max_iter) + " iterations");
  return st;
}

int ltb_set_kpe_orbital_moments(ltb_ctx* c, int on) {
  if (!c) return LTB_ERR_ARG;
  c->keep_orbital = on != 0;
  return LTB_OK;
}

int ltb_set_kpe_light_cone(ltb_ctx* c, int on) {
  if (!c) return LTB_ERR_ARG;
  c->light_cone = on != 0;
  return LTB_OK;
}

int ltb_density_diag(ltb_ctx* c, double* diagP) {
  int st = need_system(c);
  if (st) return st;
  if (!diagP) return set_err(c, LTB_ERR_ARG, "null output");
  if (c->momorb_M < 0 || c->momorb_M < c->fd_M || c->fd_M <= 0)
    return set_err(c, LTB_ERR_STATE, "needs ltb_set_kpe_orbital_moments(1), ltb_kpe_moments, ltb_density");
  double* d = nullptr;
  LTB_HIP(c, hipMalloc(&d, sizeof(double) * c->norb));
  hipLaunchKernelGGL(k_diagp, dim3(grid_for(c->norb, 256)), dim3(256), 0, c->stream, c->d_momorb, c->d_w,
                     c->fd_M, c->norb, d);
  LTB_HIP(c, hipMemcpyAsync(diagP, d, sizeof(double) * c->norb, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  (void)hipFree(d);
  return LTB_OK;
}

int ltb_get_orbital_moments(ltb_ctx* c, int M, int j0, int j1, double* out) {
  int st = need_system(c);
  if (st) return st;
  if (!out || j0 < 0 || j1 > c->norb || j0 >= j1) return set_err(c, LTB_ERR_ARG, "bad orbital range");
  if (c->momorb_M < 0 || c->momorb_M != M)
    return set_err(c, LTB_ERR_STATE, "needs ltb_set_kpe_orbital_moments(1) and ltb_kpe_moments with this M");
  const size_t w = sizeof(double) * (size_t)(j1 - j0);
  LTB_HIP(c, hipMemcpy2DAsync(out, w, c->d_momorb + j0, sizeof(double) * (size_t)c->norb, w, (size_t)M + 1,
                              hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_kpe_weights(ltb_ctx* c, int M, double* w) {
  int st = need_system(c);
  if (st) return st;
  if (!w || M != c->fd_M || !c->d_w) return set_err(c, LTB_ERR_STATE, "no ltb_density call with this M");
  LTB_HIP(c, hipMemcpyAsync(w, c->d_w, sizeof(double) * M, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_kpe_weights_at(ltb_ctx* c, int M, double mu_ha, double kT, int K, double* w) {
  int st = need_system(c);
  if (st) return st;
  if (!w || M < 2 || !std::isfinite(mu_ha)) return set_err(c, LTB_ERR_ARG, "bad arguments to ltb_kpe_weights_at");
  if (kT <= 0) K = 1;
  if (K < 1) return set_err(c, LTB_ERR_ARG, "K >= 1 quadrature nodes needed");
  if ((st = ensure_fermi(c, M, K))) return st;
  DevScalars h;
  if ((st = read_scalars(c, &h))) return st;
  h.mu = (mu_ha - h.b) / h.a;  // the H~ frame of the last ltb_rescale
  LTB_HIP(c, hipMemcpyAsync(c->d_sc, &h, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_nodes, dim3(grid_for(K, 256)), dim3(256), 0, c->stream, K, c->d_X);
  hipLaunchKernelGGL(k_coeffs, dim3(M), dim3(256), 0, c->stream, c->d_X, K, c->d_g, M, kT, c->d_sc, c->d_w);
  LTB_HIP(c, hipGetLastError());
  LTB_HIP(c, hipMemcpyAsync(w, c->d_w, sizeof(double) * M, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_kpe_stats(ltb_ctx* c, int* launches, double* secs, double* bytes) {
  if (!c) return LTB_ERR_ARG;
  if (launches) *launches = c->stat_launches;
  if (secs) *secs = c->stat_seconds;
  if (bytes) *bytes = c->stat_bytes;
  return LTB_OK;
}

int ltb_set_kpe_timing(ltb_ctx* c, int on) {
  if (!c) return LTB_ERR_ARG;
  c->timing = on != 0;
  return LTB_OK;
}

int ltb_set_kpe_tiling(ltb_ctx* c, const ltb_kpe_tiling* t) {
  if (!c || !t) return LTB_ERR_ARG;
  if (t->chunk_cols < 0 || t->lo_chunk_cols < 0 || t->atoms_per_wave < 0 || t->blocks_per_cu < 0 || t->schedule < 0 ||
      t->schedule > 2 || t->order < 0 || t->order > 2 ||
      (t->cols_per_lane != 0 && t->cols_per_lane != 1 && t->cols_per_lane != 2 && t->cols_per_lane != 4))
    return set_err(c, LTB_ERR_ARG, "bad tiling");
  c->tiling[0] = c->tiling[1] = *t;
  return LTB_OK;
}

int ltb_get_kpe_tiling(ltb_ctx* c, int precision, ltb_kpe_tiling* out) {
  int st = need_system(c);
  if (st) return st;
  if (!out) return set_err(c, LTB_ERR_ARG, "null output");
  if (precision != LTB_FP64 && precision != LTB_FP32_ACC64) return set_err(c, LTB_ERR_ARG, "bad precision");
  return kpe_resolved_tiling(c, precision, out);
}

int ltb_kpe_autotune(ltb_ctx* c, int precision, int trial_steps, ltb_kpe_tiling* best,
                     double* best_gbs) {
  int st = need_system(c);
  if (st) return st;
  if (precision != LTB_FP64 && precision != LTB_FP32_ACC64) return set_err(c, LTB_ERR_ARG, "bad precision");
  if (trial_steps < 1) trial_steps = 3;
  ltb_kpe_tiling& cur = c->tiling[precision == LTB_FP64 ? 1 : 0];  // only this precision's tiling is tuned
  const ltb_kpe_tiling saved = cur;
  std::vector<ltb_kpe_tiling> cand;
  const int chunks[] = {2048, 4096, 8192};
  const int vs[] = {1, 2, 4};
  for (int ch : chunks)
    for (int v : vs)
      for (int nt = 0; nt <= 1; nt++) {
        cand.push_back({ch, v, LTB_SCHED_BLOCKED, 0, 0, nt, saved.order, 0});
        for (int bpc : {2, 4, 8}) cand.push_back({ch, v, LTB_SCHED_SWEEP, bpc, 0, nt, saved.order, 0});
      }
  double best_rate = -1;
  ltb_kpe_tiling win = saved;
  for (const ltb_kpe_tiling& t : cand) {
    cur = t;
    ltb_kpe_tiling r;
    kpe_resolved_tiling(c, precision, &r);
    if (r.schedule != t.schedule) continue;  // sweep needs >= 8 column groups
    if (t.chunk_cols > 4 * c->norb) continue;
    double secs = 0, bytes = 0;
    st = kpe_time_steps(c, precision, trial_steps, &secs, &bytes);
    if (st) {
      cur = saved;
      return st;
    }
    double rate = bytes / secs;
    if (rate > best_rate) {
      best_rate = rate;
      win = t;
    }
  }
  cur = win;
  cur.lo_chunk_cols = 0;
  // the two-level SCC's low-order recursions: light cones cover less of the
  // system in narrower chunks, which pays in a short recursion; a whole
  // order-32 one is timed at the chosen width and at half of it
  ltb_kpe_tiling r;
  kpe_resolved_tiling(c, precision, &r);
  const int half = r.chunk_cols / 2;
  if (c->light_cone && r.schedule == LTB_SCHED_SWEEP && c->norb > r.chunk_cols && half % 4 == 0 &&
      half / (64 * r.cols_per_lane) >= 8) {
    double t_full = 0, t_half = 0;
    if ((st = kpe_time_recursion(c, precision, 32, r.chunk_cols, 2, &t_full)) ||
        (st = kpe_time_recursion(c, precision, 32, half, 2, &t_half))) {
      cur = saved;
      return st;
    }
    if (t_half < t_full) cur.lo_chunk_cols = half;
  }
  if (best) kpe_resolved_tiling(c, precision, best);
  if (best_gbs) *best_gbs = best_rate / 1e9;
  return LTB_OK;
}

}  // extern "C"
