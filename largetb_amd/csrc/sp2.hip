// Dense SP2 purification with hand-written MFMA GEMMs.
//
// Replaces the reference's vector-recursive SP2 (getSP2 withSP2_deltaH.cpp
// :1187-1215, which costs 2^depth KPE applications per vector) and its trace
// rule (computeAlphas :1217-1294): on an explicit matrix X,
//     X <- X^2          if 2 tr X > Ne   (alpha = -1)
//     X <- 2X - X^2     otherwise        (alpha = +1)
// starting from X0 = (emax I - H_total) / (emax - emin).  The bulk runs with
// FP32 operands on v_mfma_f32_32x32x2f32; the last iterations use FP64 on
// v_mfma_f64_16x16x4f64.
//
// GEMM: X is symmetric, so X^2 = X^T X and both operand tiles are row slabs
// X[k0:k0+16, i0:i0+128] and X[k0:k0+16, j0:j0+128] -- contiguous 512 B (FP32)
// rows.  Only the tiles on and above the diagonal are formed (nt (nt+1) / 2 of
// the nt^2, nt = N / 128); each off-diagonal tile is also written to its mirror
// position, so the iterate stays exactly symmetric and the MFMA work is
// (nt + 1) / (2 nt) of the full product.  Workgroup = 256 threads (2x2 waves),
// 128x128 output tile, K-slab 16 staged through LDS with a register prefetch of
// the next slab.  The epilogue applies the branch and accumulates tr(X^2) from
// the diagonal tiles (one FP64 atomic per wave).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <vector>

#include "ctx.hpp"

namespace {

constexpr int BM = 128;  // output tile (square)
constexpr int PAD = 4;   // LDS row padding (elements)
// K slab of 16 rows: the double-buffered tiles take 33.8 KB (FP32) / 67.6 KB
// (FP64); the FP64 slab carries 4096 MFMA cycles per wave, long enough to
// cover the next slab's global loads at 2 waves per SIMD

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
  static constexpr int TM = 32;  // MFMA tile edge
  static constexpr int TK = 2;   // K per instruction
  static constexpr int NACC = 16;
  typedef float acc_t __attribute__((ext_vector_type(16)));
  __device__ static acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
  }
  // operand lane map: row/col = lane % 32, k = lane / 32
  __device__ static int op_mn(int lane) { return lane & 31; }
  __device__ static int op_k(int lane) { return lane >> 5; }
  // accumulator e of lane: row 8*(e/4) + 4*(lane/32) + e%4, column lane%32
  __device__ static int acc_row(int lane, int e) { return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3); }
  __device__ static int acc_col(int lane) { return lane & 31; }
  static constexpr int RUN = 4;  // accumulators e..e+3 (e % 4 == 0) hold consecutive rows
};

template <>
struct Mfma<double> {
  static constexpr int TM = 16;
  static constexpr int TK = 4;
  static constexpr int NACC = 4;
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int op_mn(int lane) { return lane & 15; }
  __device__ static int op_k(int lane) { return lane >> 4; }
  // accumulator e of lane: row 4*e + lane/16, column lane%16 (checked by a lane probe)
  __device__ static int acc_row(int lane, int e) { return 4 * e + (lane >> 4); }
  __device__ static int acc_col(int lane) { return lane & 15; }
  static constexpr int RUN = 1;
};

// upper-triangle tile t (row-major over ti <= tj) -> (ti, tj); nt tiles per edge
__device__ inline void upper_tile(int t, int nt, int& ti, int& tj) {
  // row ti starts at s(ti) = ti nt - ti (ti - 1) / 2; invert, then correct the rounding
  const double b = 2.0 * nt + 1.0;
  int r = (int)((b - sqrt(b * b - 8.0 * t)) * 0.5);
  auto start = [nt](int q) { return q * nt - q * (q - 1) / 2; };
  while (r > 0 && start(r) > t) r--;
  while (r + 1 < nt && start(r + 1) <= t) r++;
  ti = r;
  tj = r + (t - start(r));
}

// Y = X^2 (square = 1) or 2X - X^2 (square = 0); trC += tr(X^2).
// N is a multiple of 128 (zero padding); ld = N; one workgroup per tile on or
// above the diagonal (grid = nt (nt + 1) / 2).
template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 3 : 2) void k_sp2_gemm(const T* __restrict__ X, T* __restrict__ Y, int N,
                                                  int square, double* __restrict__ trC) {
  using MF = Mfma<T>;
  constexpr int BK = 16;
  constexpr int TPW = 64 / MF::TM;                  // MFMA tiles per wave edge
  constexpr int EPT = BK * BM / 256;                // slab elements per thread
  constexpr int VEC = 16 / sizeof(T);               // elements per 16-byte load
  constexpr int NV = EPT / VEC;                     // 16-byte loads per thread per slab
  __shared__ T sA[2][BK][BM + PAD];
  __shared__ T sB[2][BK][BM + PAD];
  int ti, tj;
  upper_tile(blockIdx.x, N / BM, ti, tj);
  const int i0 = ti * BM, j0 = tj * BM;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  // slab load map: thread -> (row r, columns c..c+EPT)
  const int lr = tid / (BM / EPT), lc = (tid % (BM / EPT)) * EPT;

  typename MF::acc_t acc[TPW][TPW];
#pragma unroll
  for (int a = 0; a < TPW; a++)
#pragma unroll
    for (int b = 0; b < TPW; b++)
#pragma unroll
      for (int e = 0; e < MF::NACC; e++) acc[a][b][e] = T(0);

  typedef T vec_t __attribute__((ext_vector_type(VEC)));
  vec_t ra[NV], rb[NV];
  auto gload = [&](int k0) {
#pragma unroll
    for (int v = 0; v < NV; v++) {
      ra[v] = *reinterpret_cast<const vec_t*>(X + (size_t)(k0 + lr) * N + i0 + lc + v * VEC);
      rb[v] = *reinterpret_cast<const vec_t*>(X + (size_t)(k0 + lr) * N + j0 + lc + v * VEC);
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
      for (int q = 0; q < VEC; q++) {
        sA[buf][lr][lc + v * VEC + q] = ra[v][q];
        sB[buf][lr][lc + v * VEC + q] = rb[v][q];
      }
  };

  gload(0);
  sstore(0);
  __syncthreads();
  const int nk = N / BK;
  for (int ks = 0; ks < nk; ks++) {
    const int buf = ks & 1;
    if (ks + 1 < nk) gload((ks + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += MF::TK) {
      const int k = kk + MF::op_k(lane);
      T av[TPW], bv[TPW];
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        av[t] = sA[buf][k][wr * 64 + t * MF::TM + MF::op_mn(lane)];
        bv[t] = sB[buf][k][wc * 64 + t * MF::TM + MF::op_mn(lane)];
      }
#pragma unroll
      for (int a = 0; a < TPW; a++)
#pragma unroll
        for (int b = 0; b < TPW; b++) acc[a][b] = MF::mma(av[a], bv[b], acc[a][b]);
    }
    if (ks + 1 < nk) sstore(buf ^ 1);
    __syncthreads();
  }

  double dsum = 0;
  const bool mirror = ti != tj;
  typedef T run_t __attribute__((ext_vector_type(MF::RUN)));
#pragma unroll
  for (int a = 0; a < TPW; a++)
#pragma unroll
    for (int b = 0; b < TPW; b++) {
      const int j = j0 + wc * 64 + b * MF::TM + MF::acc_col(lane);
#pragma unroll
      for (int e = 0; e < MF::NACC; e++) {
        const int i = i0 + wr * 64 + a * MF::TM + MF::acc_row(lane, e);
        const T c = acc[a][b][e];
        const size_t o = (size_t)i * N + j;
        Y[o] = square ? c : T(2) * X[o] - c;
        if (i == j) dsum += (double)c;
      }
      if (mirror) {
        // C[j, i] = C[i, j]: rows acc_row(e .. e+RUN-1) are consecutive, so each
        // lane writes RUN consecutive elements of row j of Y
#pragma unroll
        for (int e = 0; e < MF::NACC; e += MF::RUN) {
          const int i = i0 + wr * 64 + a * MF::TM + MF::acc_row(lane, e);
          const size_t o = (size_t)j * N + i;
          run_t c;
#pragma unroll
          for (int q = 0; q < MF::RUN; q++) c[q] = acc[a][b][e + q];
          if (!square) c = T(2) * *reinterpret_cast<const run_t*>(X + o) - c;
          *reinterpret_cast<run_t*>(Y + o) = c;
        }
      }
    }
  if (ti == tj && wr == wc) {
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off);
    if (lane == 0) atomicAdd(trC, dsum);
  }
}

// X0 = (emax I - H_total) / (emax - emin), dense N x N (padding stays zero)
template <typename T>
__global__ void k_sp2_init(const int* __restrict__ blkrow, const int* __restrict__ col,
                           const double* __restrict__ h0, const double* __restrict__ v, int nnzb,
                           double emin, double emax, int N, T* __restrict__ X) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nnzb * 16) return;
  const int k = idx >> 4, e = idx & 15, r = e >> 2, s = e & 3;
  const int A = blkrow[k], B = col[k];
  double h = h0[idx];
  const bool diag = (A == B && r == s);
  if (diag) h += v[A];
  double x = ((diag ? emax : 0.0) - h) / (emax - emin);
  X[(size_t)(4 * A + r) * N + 4 * B + s] = (T)x;
}

// tr(X^2) = sum_ij X_ij^2 for symmetric X (block partials, FP64)
template <typename T>
__global__ __launch_bounds__(256) void k_frob2(const T* __restrict__ X, size_t n, double* __restrict__ part) {
  double s = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double x = (double)X[i];
    s += x * x;
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_f32_to_f64(const float* __restrict__ a, double* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) b[i] = (double)a[i];
}

template <typename T>
__global__ void k_sp2_diag(const T* __restrict__ X, int N, int norb, double* __restrict__ d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < norb) d[i] = (double)X[(size_t)i * N + i];
}

}  // namespace

int sp2_dense_impl(ltb_ctx* c, double Ne, int max_iter, int fp64_tail, double* diagP, double* pop,
                   ltb_sp2_result* res) {
  const int norb = c->norb;
  const int N = (norb + BM - 1) / BM * BM;
  const size_t n2 = (size_t)N * N;
  // Gershgorin bounds of H_total (v from the last ltb_set_populations)
  double emin, emax;
  int st = ltb_gershgorin(c, &emin, &emax);
  if (st) return st;
  int st2;
  const size_t need = n2 * (fp64_tail > 0 ? sizeof(double) : sizeof(float));
  if (c->sbytes < need) {  // both ping-pong buffers always share one capacity
    (void)hipFree(c->d_sa);
    (void)hipFree(c->d_sb);
    c->d_sa = c->d_sb = nullptr;
    c->sbytes = 0;
    if (hipMalloc(&c->d_sa, need) != hipSuccess || hipMalloc(&c->d_sb, need) != hipSuccess) {
      (void)hipFree(c->d_sa);
      c->d_sa = nullptr;
      return set_err(c, LTB_ERR_NOMEM, "SP2 needs 2 x " + std::to_string(need) + " bytes");
    }
    c->sbytes = need;
  }
  double* d_tr = c->d_red;  // scratch scalar
  hipEvent_t e0, e1;
  LTB_HIP(c, hipEventCreate(&e0));
  LTB_HIP(c, hipEventCreate(&e1));
  double gemm_s = 0, flops = 0;
  const double nocc = Ne / 2.0;
  // trace of X0 on the host from the diagonal
  double trH = 0;
  {
    std::vector<double> d0(norb), v(c->nat);
    LTB_HIP(c, hipMemcpy(d0.data(), c->d_diag0, sizeof(double) * norb, hipMemcpyDeviceToHost));
    LTB_HIP(c, hipMemcpy(v.data(), c->d_v, sizeof(double) * c->nat, hipMemcpyDeviceToHost));
    for (int i = 0; i < norb; i++) trH += d0[i] + v[i / 4];
  }
  double trX = (emax * norb - trH) / (emax - emin);
  const int nt = N / BM;
  dim3 grid(nt * (nt + 1) / 2), block(256);
  // branch rule: the reference's (square iff 2 tr X > Ne, computeAlphas :1277-1281)
  // in the bulk; in the FP64 tail after an FP32 bulk, Niklasson's rule (square
  // iff |tr X^2 - Nocc| < |2 tr X - tr X^2 - Nocc|), which also contracts
  // eigenvalues that FP32 rounding left slightly outside [0, 1].
  auto run_phase = [&](auto tag, int max_it, bool stall_stop, bool nearest_rule, int* iters,
                       double* idem_out) -> int {
    using T = decltype(tag);
    T* A = (T*)c->d_sa;
    T* B = (T*)c->d_sb;
    double idem_hist[3] = {INFINITY, INFINITY, INFINITY};
    int it = 0;
    for (; it < max_it; it++) {
      int square = 2.0 * trX > Ne ? 1 : 0;
      if (nearest_rule) {
        const int nb = 1024;
        hipLaunchKernelGGL((k_frob2<T>), dim3(nb), dim3(256), 0, c->stream, (const T*)A, n2, c->d_red + 8);
        std::vector<double> part(nb);
        LTB_HIP(c, hipMemcpyAsync(part.data(), c->d_red + 8, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
        LTB_HIP(c, hipStreamSynchronize(c->stream));
        double t2 = 0;
        for (double x : part) t2 += x;
        square = std::fabs(t2 - nocc) < std::fabs(2.0 * trX - t2 - nocc) ? 1 : 0;
      }
      LTB_HIP(c, hipMemsetAsync(d_tr, 0, sizeof(double), c->stream));
      LTB_HIP(c, hipEventRecord(e0, c->stream));
      hipLaunchKernelGGL((k_sp2_gemm<T>), grid, block, 0, c->stream, A, B, N, square, d_tr);
      LTB_HIP(c, hipEventRecord(e1, c->stream));
      LTB_HIP(c, hipGetLastError());
      double trC = 0;
      LTB_HIP(c, hipMemcpyAsync(&trC, d_tr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
      LTB_HIP(c, hipStreamSynchronize(c->stream));
      float ms = 0;
      LTB_HIP(c, hipEventElapsedTime(&ms, e0, e1));
      gemm_s += ms * 1e-3;
      flops += 2.0 * BM * BM * (double)N * (double)grid.x;  // MFMA work of the formed tiles
      const double idem = std::fabs(trX - trC);
      trX = square ? trC : 2.0 * trX - trC;
      std::swap(A, B);
      c->d_sa = A;
      c->d_sb = B;
      idem_hist[0] = idem_hist[1];
      idem_hist[1] = idem_hist[2];
      idem_hist[2] = idem;
      *idem_out = idem;
      if (!std::isfinite(idem)) return set_err(c, LTB_ERR_NONFINITE, "non-finite trace in SP2");
      // stop when the idempotency error no longer falls (rounding floor)
      // (quadratic regime reached and no halving over two steps)
      if (stall_stop && it >= 4 && idem < 1e-4 * norb && idem >= 0.5 * idem_hist[0]) {
        it++;
        break;
      }
      if (idem < 1e-12 * norb) {
        it++;
        break;
      }
    }
    *iters = it;
    return LTB_OK;
  };
  int it32 = 0, it64 = 0;
  double idem = 0;
  const int ginit = (c->nnzb * 16 + 255) / 256;
  if (max_iter > 0) {
    // FP32 bulk until the idempotency error stalls, then widen for the FP64 tail
    LTB_HIP(c, hipMemsetAsync(c->d_sa, 0, n2 * sizeof(float), c->stream));
    hipLaunchKernelGGL((k_sp2_init<float>), dim3(ginit), dim3(256), 0, c->stream, c->d_blkrow,
                       c->d_col, c->d_h0, c->d_v, c->nnzb, emin, emax, N, (float*)c->d_sa);
    LTB_HIP(c, hipGetLastError());
    if ((st2 = run_phase(float(0), max_iter, true, false, &it32, &idem))) return st2;
    if (fp64_tail > 0) {
      hipLaunchKernelGGL(k_f32_to_f64, dim3(4096), dim3(256), 0, c->stream, (const float*)c->d_sa,
                         (double*)c->d_sb, n2);
      LTB_HIP(c, hipGetLastError());
      std::swap(c->d_sa, c->d_sb);
      if ((st2 = run_phase(double(0), fp64_tail, false, true, &it64, &idem))) return st2;
    }
  } else {
    // FP64 throughout (the reference's precision), fp64_tail is the iteration cap
    LTB_HIP(c, hipMemsetAsync(c->d_sa, 0, n2 * sizeof(double), c->stream));
    hipLaunchKernelGGL((k_sp2_init<double>), dim3(ginit), dim3(256), 0, c->stream, c->d_blkrow,
                       c->d_col, c->d_h0, c->d_v, c->nnzb, emin, emax, N, (double*)c->d_sa);
    LTB_HIP(c, hipGetLastError());
    if ((st2 = run_phase(double(0), fp64_tail, true, false, &it64, &idem))) return st2;
  }
  // idempotency of the final iterate: |tr X - tr X^2|, tr X^2 = ||X||_F^2
  {
    const int nb = 1024;
    const bool f64 = fp64_tail > 0 || max_iter == 0;
    if (f64)
      hipLaunchKernelGGL((k_frob2<double>), dim3(nb), dim3(256), 0, c->stream, (const double*)c->d_sa, n2, c->d_red + 8);
    else
      hipLaunchKernelGGL((k_frob2<float>), dim3(nb), dim3(256), 0, c->stream, (const float*)c->d_sa, n2, c->d_red + 8);
    std::vector<double> part(nb);
    LTB_HIP(c, hipMemcpyAsync(part.data(), c->d_red + 8, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    LTB_HIP(c, hipStreamSynchronize(c->stream));
    double t2 = 0;
    for (double x : part) t2 += x;
    idem = std::fabs(trX - t2);
  }
  std::vector<double> d(norb);
  double* dd = nullptr;
  LTB_HIP(c, hipMalloc(&dd, sizeof(double) * norb));
  if (fp64_tail > 0 || max_iter == 0)
    hipLaunchKernelGGL((k_sp2_diag<double>), dim3((norb + 255) / 256), dim3(256), 0, c->stream,
                       (const double*)c->d_sa, N, norb, dd);
  else
    hipLaunchKernelGGL((k_sp2_diag<float>), dim3((norb + 255) / 256), dim3(256), 0, c->stream,
                       (const float*)c->d_sa, N, norb, dd);
  LTB_HIP(c, hipMemcpyAsync(d.data(), dd, sizeof(double) * norb, hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  (void)hipFree(dd);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (diagP) std::copy(d.begin(), d.end(), diagP);
  if (pop)
    for (int a = 0; a < c->nat; a++) pop[a] = d[4 * a] + d[4 * a + 1] + d[4 * a + 2] + d[4 * a + 3];
  if (res) {
    res->iterations = it32 + it64;
    res->fp64_iterations = it64;
    double tr = 0;
    for (double x : d) tr += x;
    res->trace = tr;
    res->idem = idem;
    res->gemm_seconds = gemm_s;
    res->gemm_flops = flops;
  }
  return LTB_OK;
}

extern "C" int ltb_sp2_dense(ltb_ctx* c, double Ne, int max_iter, int fp64_tail, double* diagP,
                             double* pop, ltb_sp2_result* res) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  if (!(Ne > 0) || max_iter < 0 || fp64_tail < 0 || max_iter > 1000 || fp64_tail > 1000 ||
      (max_iter == 0 && fp64_tail == 0))
    return set_err(c, LTB_ERR_ARG, "bad SP2 arguments");
  return sp2_dense_impl(c, Ne, max_iter, fp64_tail, diagP, pop, res);
}
