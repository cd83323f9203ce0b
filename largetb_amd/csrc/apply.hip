// getKPE (withSP2_deltaH.cpp:1068-1185) on a block of vectors:
//     Y = w_0 X + sum_{m=1}^{M-1} w_m T_m(H~) X,   w = (c_0/2, c_m g_m)
// The block is norb x ld, row-major (columns contiguous), as in cheb.hip.  One
// launch advances the recursion one step and accumulates theta += w_m t_m in
// the same pass (the reference's MatMult + 6 BLAS-1 calls per step).
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ctx.hpp"

namespace {

// MODE 0: t1 = H t0, theta = w0 t0 + w1 t1 ; MODE 2: t_{n+1} = 2 H t_n - t_{n-1}, theta += w t
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_apply_step(const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const T* __restrict__ hv,
                                                    const T* __restrict__ xcur,
                                                    T* __restrict__ xnext,
                                                    double* __restrict__ theta, int nat, int ld,
                                                    double w0, double w1) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int A = blockIdx.y * 4 + wave;
  const int c = blockIdx.x * 64 + lane;
  if (A >= nat) return;
  T acc[4] = {0, 0, 0, 0};
  for (int k = rowptr[A]; k < rowptr[A + 1]; k++) {
    const int B = col[k];
    const T* h = hv + 16 * (size_t)k;
    T x[4];
    for (int s = 0; s < 4; s++) x[s] = xcur[(size_t)(4 * B + s) * ld + c];
    for (int r = 0; r < 4; r++)
      for (int s = 0; s < 4; s++) acc[r] = fma(h[4 * r + s], x[s], acc[r]);
  }
  for (int r = 0; r < 4; r++) {
    const size_t o = (size_t)(4 * A + r) * ld + c;
    T y;
    if (MODE == 0) {
      y = acc[r];
      theta[o] = w0 * (double)xcur[o] + w1 * (double)y;
    } else {
      y = T(2) * acc[r] - xnext[o];
      theta[o] += w1 * (double)y;
    }
    xnext[o] = y;
  }
}

// the recursion on device buffers: a holds t_0 (overwritten), b is scratch,
// th receives theta; all norb x ld row-major
template <typename T>
int apply_core(ltb_ctx* c, int M, const double* w, int ld, T* a, T* b, double* th) {
  const T* hv = sizeof(T) == 4 ? (const T*)c->d_ht32 : (const T*)c->d_ht64;
  dim3 grid(ld / 64, (c->nat + 3) / 4), block(256);
  // step 1: t1 = H t0 into b; theta = w0 t0 + w1 t1
  hipLaunchKernelGGL((k_apply_step<T, 0>), grid, block, 0, c->stream, c->d_rowptr, c->d_col, hv, a,
                     b, th, c->nat, ld, w[0], M > 1 ? w[1] : 0.0);
  // steps m >= 2: cur alternates b, a; next overwrites the older one
  for (int m = 2; m < M; m++) {
    T* cur = (m & 1) ? a : b;
    T* nxt = (m & 1) ? b : a;
    hipLaunchKernelGGL((k_apply_step<T, 2>), grid, block, 0, c->stream, c->d_rowptr, c->d_col, hv,
                       cur, nxt, th, c->nat, ld, 0.0, w[m]);
  }
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

// device buffers of one call, freed on every return path once the stream has
// drained (hipFree waits for the device)
struct CallBuffers {
  void* p[3] = {nullptr, nullptr, nullptr};
  ~CallBuffers() {
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

template <typename T>
int run_apply(ltb_ctx* c, int M, const double* w, int ncols, const double* X, double* Y) {
  const int ld = (ncols + 63) / 64 * 64;
  const size_t n = (size_t)c->norb * ld;
  std::vector<T> hx(n, T(0));
  for (int j = 0; j < ncols; j++)
    for (int r = 0; r < c->norb; r++) hx[(size_t)r * ld + j] = (T)X[(size_t)j * c->norb + r];
  CallBuffers buf;
  LTB_HIP(c, hipMalloc(&buf.p[0], n * sizeof(T)));
  LTB_HIP(c, hipMalloc(&buf.p[1], n * sizeof(T)));
  LTB_HIP(c, hipMalloc(&buf.p[2], n * sizeof(double)));
  T* a = (T*)buf.p[0];
  T* b = (T*)buf.p[1];
  double* th = (double*)buf.p[2];
  LTB_HIP(c, hipMemcpyAsync(a, hx.data(), n * sizeof(T), hipMemcpyHostToDevice, c->stream));
  int st = apply_core<T>(c, M, w, ld, a, b, th);
  if (st) {
    (void)hipStreamSynchronize(c->stream);
    return st;
  }
  std::vector<double> ht(n);
  LTB_HIP(c, hipMemcpyAsync(ht.data(), th, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < ncols; j++)
    for (int r = 0; r < c->norb; r++) Y[(size_t)j * c->norb + r] = ht[(size_t)r * ld + j];
  return LTB_OK;
}

// ---- allocation-free variant ------------------------------------------------------
// Its recursion vectors t_n, t_{n-1} and theta live in one file-scope device
// array (each GPU holds its own copy for the life of the process).  Host code
// never names the array except to look up its device address once per
// context, in noalloc_scratch_init below; the kernels get that address.
constexpr size_t kScratchBytes = size_t(96) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t align256(size_t n) { return (n + 255) / 256 * 256; }

// X: [ncols][norb] FP64 columns -> a: [norb][ld] row-major in T, zero for the
// padding columns ncols <= j < ld (the array keeps the last call's contents)
template <typename T>
__global__ __launch_bounds__(256) void k_cols_to_block(const double* __restrict__ X, int ncols, int norb,
                                                       int ld, T* __restrict__ a) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)norb * ld) return;
  const int r = (int)(i / ld), j = (int)(i - (size_t)r * ld);
  a[i] = j < ncols ? (T)X[(size_t)j * norb + r] : T(0);
}

// theta: [norb][ld] -> Y: [ncols][norb]
__global__ __launch_bounds__(256) void k_block_to_cols(const double* __restrict__ th, int ncols, int norb,
                                                       int ld, double* __restrict__ Y) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)ncols * norb) return;
  const int j = (int)(i / norb), r = (int)(i - (size_t)j * norb);
  Y[i] = th[(size_t)r * ld + j];
}

template <typename T>
int run_apply_noalloc(ltb_ctx* c, int M, const double* w, int ncols, const double* dX, double* dY) {
  const int ld = (ncols + 63) / 64 * 64;
  const size_t n = (size_t)c->norb * ld;
  const size_t ob = align256(n * sizeof(T)), oth = 2 * ob, need = oth + n * sizeof(double);
  if (need > kScratchBytes) {
    const size_t per64 = 2 * align256((size_t)c->norb * 64 * sizeof(T)) + (size_t)c->norb * 64 * sizeof(double);
    return set_err(c, LTB_ERR_ARG,
                   "ltb_kpe_apply_noalloc: " + std::to_string(ncols) + " columns of " + std::to_string(c->norb) +
                       " orbitals need " + std::to_string(need) + " bytes of scratch; the maximum is " +
                       std::to_string(kScratchBytes) + " (at most " + std::to_string(64 * (kScratchBytes / per64)) +
                       " columns per call in this precision)");
  }
  unsigned char* base = static_cast<unsigned char*>(c->d_scratch);
  T* a = reinterpret_cast<T*>(base);
  T* b = reinterpret_cast<T*>(base + ob);
  double* th = reinterpret_cast<double*>(base + oth);
  hipLaunchKernelGGL((k_cols_to_block<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, dX, ncols,
                     c->norb, ld, a);
  LTB_HIP(c, hipGetLastError());
  int st = apply_core<T>(c, M, w, ld, a, b, th);
  if (st) {
    (void)hipStreamSynchronize(c->stream);
    return st;
  }
  const size_t ny = (size_t)ncols * c->norb;
  hipLaunchKernelGGL(k_block_to_cols, dim3((unsigned)((ny + 255) / 256)), dim3(256), 0, c->stream, th, ncols,
                     c->norb, ld, dY);
  LTB_HIP(c, hipGetLastError());
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

}  // namespace

// device address of this GPU's scratch array, looked up once per context with
// the context's device current (called by ltb_create)
int noalloc_scratch_init(ltb_ctx* c) {
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return set_err(c, LTB_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  c->d_scratch = p;
  return LTB_OK;
}

int kpe_apply_device(ltb_ctx* c, int M, const double* w, int ld, void* a, void* b, double* theta, int precision) {
  if (precision == LTB_FP32_ACC64) return apply_core<float>(c, M, w, ld, (float*)a, (float*)b, theta);
  return apply_core<double>(c, M, w, ld, (double*)a, (double*)b, theta);
}

extern "C" int ltb_kpe_apply(ltb_ctx* c, int M, const double* w, int ncols, const double* X,
                             double* Y, int precision) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  if (M < 1 || !w || ncols < 1 || !X || !Y) return set_err(c, LTB_ERR_ARG, "bad kpe_apply arguments");
  if (precision == LTB_FP32_ACC64) return run_apply<float>(c, M, w, ncols, X, Y);
  if (precision == LTB_FP64) return run_apply<double>(c, M, w, ncols, X, Y);
  return set_err(c, LTB_ERR_ARG, "bad precision");
}

extern "C" int ltb_kpe_apply_noalloc(ltb_ctx* c, int M, const double* w, int ncols, const double* dX, double* dY,
                                     int precision) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  if (M < 1 || !w || ncols < 1 || !dX || !dY) return set_err(c, LTB_ERR_ARG, "bad kpe_apply_noalloc arguments");
  if (!c->d_scratch) return set_err(c, LTB_ERR_STATE, "no scratch address for this context's device");
  LTB_HIP(c, hipSetDevice(c->device));
  if (precision == LTB_FP32_ACC64) return run_apply_noalloc<float>(c, M, w, ncols, dX, dY);
  if (precision == LTB_FP64) return run_apply_noalloc<double>(c, M, w, ncols, dX, dY);
  return set_err(c, LTB_ERR_ARG, "bad precision");
}
