// Exact-trace Chebyshev moments of H~ by a block SpMM over 4x4 BSR blocks.
//
// Replaces the reference's per-vector getKPE recursion (withSP2_deltaH.cpp
// :1139-1168, MatMult + 6 BLAS-1 calls per step) applied to every unit vector
// (the commented exact mode :1361-1371).  A chunk of `ld` unit vectors
// e_{j0..j0+ld-1} is advanced together: T = [t_n] is norb x ld, row-major
// (orbital-major, columns contiguous), and one launch performs
//     t_{n+1} = 2 H~ t_n - t_{n-1}      (t_1 = H~ t_0, t_0 = identity columns)
// for all columns, in place over t_{n-1}.  The same pass accumulates, per
// column j, the dots <t_{n+1}, t_n> and <t_{n+1}, t_{n+1}>; by
//     T_{2n+1} = 2 T_{n+1} T_n - T_1,   T_{2n+2} = 2 T_{n+1}^2 - T_0
// they give [T_{2n+1}]_jj and [T_{2n+2}]_jj, so M moments take M/2 launches.
//
// Work mapping (CDNA4, wave64): a wave owns one atom (block row) at a time
// and 64*V columns; every lane holds V consecutive columns, so each row read
// or write of a wave is one contiguous 64*V*sizeof(T) segment.  The atom's
// block-row pointers, column indices and 4x4 H~ blocks are wave-uniform and
// come through scalar loads.  A workgroup is 4 waves (4*apw atoms, same
// columns); its per-column dot partials are summed in LDS and written once
// per workgroup, then folded by k_cheb_reduce.  Workgroup ids are remapped so
// that each XCD sweeps a contiguous range of (column group, atom range).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "ctx.hpp"

namespace {

template <typename T, int V>
struct VecT;
template <>
struct VecT<float, 4> {
  using type = float4;
};
template <>
struct VecT<float, 2> {
  using type = float2;
};
template <>
struct VecT<float, 1> {
  using type = float;
};
template <>
struct VecT<double, 2> {
  using type = double2;
};
template <>
struct VecT<double, 1> {
  using type = double;
};

template <typename T, int V>
__device__ __forceinline__ void load_row(const T* p, T (&x)[V]) {
  using W = typename VecT<T, V>::type;
  W w = *reinterpret_cast<const W*>(p);
  __builtin_memcpy(x, &w, sizeof(W));
}

template <typename T, int V>
__device__ __forceinline__ void store_row(T* p, const T (&x)[V]) {
  using W = typename VecT<T, V>::type;
  W w;
  __builtin_memcpy(&w, x, sizeof(W));
  *reinterpret_cast<W*>(p) = w;
}

// MODE 0: t_1 = H~ t_0 (t_0 implicit identity columns, nothing read)
// MODE 1: t_2 = 2 H~ t_1 - t_0 (t_0 implicit)
// MODE 2: t_{n+1} = 2 H~ t_n - t_{n-1}, t_{n-1} read from and overwritten in xnext
template <typename T, int V, int MODE>
__global__ __launch_bounds__(256) void k_cheb_step(const int* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const T* __restrict__ hv,
                                                   const T* __restrict__ xcur,
                                                   T* __restrict__ xnext, int nat, int ld, int j0,
                                                   int nvalid, int apw, int n_sg,
                                                   double* __restrict__ part) {
  constexpr int W = 64 * V;
  int bid = blockIdx.x;
  const int nb = gridDim.x;
  if ((nb & 7) == 0) bid = (bid & 7) * (nb >> 3) + (bid >> 3);  // XCD-contiguous ranges
  const int cg = bid / n_sg;
  const int sg = bid - cg * n_sg;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int c0 = cg * W + lane * V;
  const int a_begin = (sg * 4 + wave) * apw;
  const int a_end = min(a_begin + apw, nat);

  T sxy[V], syy[V];
#pragma unroll
  for (int v = 0; v < V; v++) sxy[v] = syy[v] = T(0);

  for (int A = a_begin; A < a_end; ++A) {
    const int kb = rowptr[A], ke = rowptr[A + 1];
    T acc[4][V], xa[4][V];
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int v = 0; v < V; v++) acc[r][v] = T(0);
    for (int k = kb; k < ke; ++k) {
      const int B = col[k];
      const T* h = hv + 16 * (size_t)k;
      T x[4][V];
      if (MODE == 0) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int v = 0; v < V; v++)
            x[s][v] = (4 * B + s == j0 + c0 + v && c0 + v < nvalid) ? T(1) : T(0);
      } else {
#pragma unroll
        for (int s = 0; s < 4; s++) load_row<T, V>(xcur + (size_t)(4 * B + s) * ld + c0, x[s]);
      }
      if (k == kb) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
          for (int v = 0; v < V; v++) xa[s][v] = x[s][v];
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const T hrs = h[4 * r + s];
#pragma unroll
          for (int v = 0; v < V; v++) acc[r][v] = fma(hrs, x[s][v], acc[r][v]);
        }
    }
    T* out = xnext + (size_t)(4 * A) * ld + c0;
    T y[4][V];
    if (MODE == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int v = 0; v < V; v++) y[r][v] = acc[r][v];
    } else if (MODE == 1) {
#pragma unroll
      for (int r = 0; r < 4; r++)
#pragma unroll
        for (int v = 0; v < V; v++)
          y[r][v] = T(2) * acc[r][v] - ((4 * A + r == j0 + c0 + v && c0 + v < nvalid) ? T(1) : T(0));
    } else {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        T p[V];
        load_row<T, V>(out + (size_t)r * ld, p);
#pragma unroll
        for (int v = 0; v < V; v++) y[r][v] = T(2) * acc[r][v] - p[v];
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) store_row<T, V>(out + (size_t)r * ld, y[r]);
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        sxy[v] = fma(y[r][v], xa[r][v], sxy[v]);
        syy[v] = fma(y[r][v], y[r][v], syy[v]);
      }
    }
  }

  __shared__ double sh[4][2][W];
#pragma unroll
  for (int v = 0; v < V; v++) {
    sh[wave][0][lane * V + v] = (double)sxy[v];
    sh[wave][1][lane * V + v] = (double)syy[v];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * W; e += 256) {
    const int which = e / W, cc = e - which * W;
    double s = sh[0][which][cc] + sh[1][which][cc] + sh[2][which][cc] + sh[3][which][cc];
    part[((size_t)sg * 2 + which) * ld + cg * W + cc] = s;
  }
}

// Fold the workgroup partials of step n into per-atom moments:
//   n == 0: mu_1 = <t_1, t_0>,           mu_2 = 2 <t_1, t_1> - 1
//   n >= 1: mu_{2n+1} = 2 <t_{n+1}, t_n> - mu_1,  mu_{2n+2} = 2 |t_{n+1}|^2 - 1
// and sum the 4 columns (orbitals) of each atom.  mom is [M+1][nat].
__global__ __launch_bounds__(256) void k_cheb_reduce(const double* __restrict__ part, int n_sg,
                                                     int ld, int nvalid, int j0, int n, int M,
                                                     double* __restrict__ mu1col,
                                                     double* __restrict__ mom, int nat) {
  const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  double sxy = 0, syy = 0;
  for (int s = slice; s < n_sg; s += 4) {
    sxy += part[(size_t)(2 * s) * ld + c];
    syy += part[(size_t)(2 * s + 1) * ld + c];
  }
  __shared__ double sh[2][4][64];
  sh[0][slice][lane] = sxy;
  sh[1][slice][lane] = syy;
  __syncthreads();
  if (slice != 0) return;
  sxy = (sh[0][0][lane] + sh[0][1][lane]) + (sh[0][2][lane] + sh[0][3][lane]);
  syy = (sh[1][0][lane] + sh[1][1][lane]) + (sh[1][2][lane] + sh[1][3][lane]);
  const bool valid = c < nvalid;
  double modd;
  if (n == 0) {
    modd = sxy;
    mu1col[c] = sxy;
  } else {
    modd = 2.0 * sxy - mu1col[c];
  }
  double meven = 2.0 * syy - (valid ? 1.0 : 0.0);
  modd += __shfl_xor(modd, 1);
  modd += __shfl_xor(modd, 2);
  meven += __shfl_xor(meven, 1);
  meven += __shfl_xor(meven, 2);
  if ((lane & 3) == 0 && valid) {
    const int atom = (j0 + c) >> 2;
    if (n == 0) mom[atom] = 4.0;
    const int mo = 2 * n + 1, me = 2 * n + 2;
    if (mo <= M) mom[(size_t)mo * nat + atom] = modd;
    if (me <= M) mom[(size_t)me * nat + atom] = meven;
  }
}

__global__ __launch_bounds__(256) void k_mom_total(const double* __restrict__ mom, int nat,
                                                   double* __restrict__ tot) {
  const int m = blockIdx.x;
  double s = 0;
  for (int a = threadIdx.x; a < nat; a += 256) s += mom[(size_t)m * nat + a];
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[m] = sh[0];
}

struct Plan {
  int V, W, ld, apw, n_sg;
};

template <typename T>
Plan make_plan(const ltb_ctx* c) {
  Plan p;
  const bool big = c->norb >= 4096;
  if (sizeof(T) == 4)
    p.V = big ? 4 : 1;
  else
    p.V = big ? 2 : 1;
  p.W = 64 * p.V;
  int want = c->chunk_cols > 0 ? c->chunk_cols : (sizeof(T) == 4 ? 4096 : 2048);
  int ld = std::min(want, c->norb);
  p.ld = (ld + p.W - 1) / p.W * p.W;
  int n_cg = p.ld / p.W;
  if (c->atoms_per_wave > 0) {
    p.apw = c->atoms_per_wave;
  } else {
    // enough workgroups to fill the chip ~8 deep, but few partials per column
    int target_blocks = 8 * std::max(c->cu_count, 1);
    int sg_want = std::max(1, (target_blocks + n_cg - 1) / n_cg);
    p.apw = std::max(1, (c->nat + 4 * sg_want - 1) / (4 * sg_want));
  }
  p.n_sg = (c->nat + 4 * p.apw - 1) / (4 * p.apw);
  return p;
}

template <typename T, int V, int MODE>
void launch_step(const ltb_ctx* c, const T* hv, const T* xcur, T* xnext, int ld, int j0, int nvalid,
                 const Plan& p) {
  const int n_cg = ld / (64 * V);
  dim3 grid(n_cg * p.n_sg), block(256);
  hipLaunchKernelGGL((k_cheb_step<T, V, MODE>), grid, block, 0, c->stream, c->d_rowptr, c->d_col,
                     hv, xcur, xnext, c->nat, ld, j0, nvalid, p.apw, p.n_sg, c->d_part);
}

template <typename T, int V>
void launch_mode(const ltb_ctx* c, int mode, const T* hv, const T* xcur, T* xnext, int ld, int j0,
                 int nvalid, const Plan& p) {
  if (mode == 0)
    launch_step<T, V, 0>(c, hv, xcur, xnext, ld, j0, nvalid, p);
  else if (mode == 1)
    launch_step<T, V, 1>(c, hv, xcur, xnext, ld, j0, nvalid, p);
  else
    launch_step<T, V, 2>(c, hv, xcur, xnext, ld, j0, nvalid, p);
}

template <typename T>
int run_moments(ltb_ctx* c, int M) {
  const Plan p = make_plan<T>(c);
  const int nsteps = (M + 1) / 2;  // step n yields moments up to 2n+2
  size_t xb = (size_t)c->norb * p.ld * sizeof(T);
  int st;
  if ((st = ensure_buffer(c, &c->d_xa, &c->xabytes, xb))) return st;
  if ((st = ensure_buffer(c, &c->d_xb, &c->xbbytes, xb))) return st;
  if ((st = ensure_buffer(c, (void**)&c->d_part, &c->partbytes,
                          (size_t)p.n_sg * 2 * p.ld * sizeof(double))))
    return st;
  if ((st = ensure_buffer(c, (void**)&c->d_mu1, &c->mu1bytes, (size_t)p.ld * sizeof(double))))
    return st;
  if ((st = ensure_buffer(c, (void**)&c->d_mom, &c->mombytes,
                          (size_t)(M + 1) * c->nat * sizeof(double) + (size_t)(M + 1) * sizeof(double))))
    return st;
  c->d_momtot = c->d_mom + (size_t)(M + 1) * c->nat;
  const T* hv = sizeof(T) == 4 ? (const T*)c->d_ht32 : (const T*)c->d_ht64;
  T* A = (T*)c->d_xa;
  T* B = (T*)c->d_xb;

  const int nlaunch_est = ((c->norb + p.ld - 1) / p.ld) * nsteps;
  if (c->timing && (int)c->events.size() < 2 * nlaunch_est) {
    while ((int)c->events.size() < 2 * nlaunch_est) {
      hipEvent_t e;
      LTB_HIP(c, hipEventCreate(&e));
      c->events.push_back(e);
    }
  }
  const double hbytes = (double)c->nnzb * (16.0 * sizeof(T) + 4.0) + 4.0 * (c->nat + 1);
  int launches = 0;
  double bytes = 0;
  for (int j0 = 0; j0 < c->norb; j0 += p.ld) {
    const int nvalid = std::min(p.ld, c->norb - j0);
    const int ldc = (nvalid + p.W - 1) / p.W * p.W;
    for (int n = 0; n < nsteps; n++) {
      const int mode = n == 0 ? 0 : (n == 1 ? 1 : 2);
      T* cur = (n & 1) ? B : A;
      T* nxt = (n == 0) ? B : ((n & 1) ? A : B);
      if (c->timing) LTB_HIP(c, hipEventRecord(c->events[2 * launches], c->stream));
      if (p.V == 4)
        launch_mode<T, (sizeof(T) == 4 ? 4 : 2)>(c, mode, hv, cur, nxt, ldc, j0, nvalid, p);
      else if (p.V == 2)
        launch_mode<T, 2>(c, mode, hv, cur, nxt, ldc, j0, nvalid, p);
      else
        launch_mode<T, 1>(c, mode, hv, cur, nxt, ldc, j0, nvalid, p);
      if (c->timing) LTB_HIP(c, hipEventRecord(c->events[2 * launches + 1], c->stream));
      LTB_HIP(c, hipGetLastError());
      const int streams = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
      bytes += (double)streams * sizeof(T) * c->norb * (double)nvalid + hbytes;
      launches++;
      hipLaunchKernelGGL(k_cheb_reduce, dim3(ldc / 64), dim3(256), 0, c->stream, c->d_part, p.n_sg,
                         ldc, nvalid, j0, n, M, c->d_mu1, c->d_mom, c->nat);
      LTB_HIP(c, hipGetLastError());
    }
  }
  hipLaunchKernelGGL(k_mom_total, dim3(M + 1), dim3(256), 0, c->stream, c->d_mom, c->nat,
                     c->d_momtot);
  LTB_HIP(c, hipGetLastError());
  c->stat_launches = launches;
  c->stat_bytes = bytes;
  c->stat_seconds = 0;
  if (c->timing) {
    LTB_HIP(c, hipStreamSynchronize(c->stream));
    double tot = 0;
    for (int i = 0; i < launches; i++) {
      float ms = 0;
      LTB_HIP(c, hipEventElapsedTime(&ms, c->events[2 * i], c->events[2 * i + 1]));
      tot += ms * 1e-3;
    }
    c->stat_seconds = tot;
  }
  c->mom_M = M;
  return LTB_OK;
}

}  // namespace

int kpe_moments_impl(ltb_ctx* c, int M, int precision) {
  if (precision == LTB_FP32_ACC64) return run_moments<float>(c, M);
  return run_moments<double>(c, M);
}
