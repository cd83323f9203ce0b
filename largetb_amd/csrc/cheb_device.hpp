// Device routines of the Chebyshev block-SpMM shared by the single-structure
// kernels (cheb.hip) and the batched kernels (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace ltbdev {

template <typename T, int V>
struct VecT {
  typedef T type __attribute__((ext_vector_type(V)));
};
template <typename T>
struct VecT<T, 1> {
  typedef T type;
};

template <typename T, int V, bool NT>
__device__ __forceinline__ void load_row(const T* p, T (&x)[V]) {
  using W = typename VecT<T, V>::type;
  W w;
  if (NT)
    w = __builtin_nontemporal_load(reinterpret_cast<const W*>(p));
  else
    w = *reinterpret_cast<const W*>(p);
  __builtin_memcpy(x, &w, sizeof(W));
}

template <typename T, int V, bool NT>
__device__ __forceinline__ void store_row(T* p, const T (&x)[V]) {
  using W = typename VecT<T, V>::type;
  W w;
  __builtin_memcpy(&w, x, sizeof(W));
  if (NT)
    __builtin_nontemporal_store(w, reinterpret_cast<W*>(p));
  else
    *reinterpret_cast<W*>(p) = w;
}

// One block row (atom A) of one recursion step for the lane's V columns.
// MODE 0: t_1 = H~ t_0 (t_0 implicit identity columns, nothing read)
// MODE 1: t_2 = 2 H~ t_1 - t_0 (t_0 implicit)
// MODE 2: t_{n+1} = 2 H~ t_n - t_{n-1}, t_{n-1} read from and overwritten in xnext
// MODE 3: t_1 = H~ t_0 with t_0 explicit vectors read from xcur
template <typename T, int V, int MODE, bool NT>
__device__ __forceinline__ void atom_step(const int* __restrict__ rowptr,
                                          const int* __restrict__ col, const T* __restrict__ hv,
                                          const T* __restrict__ xcur, T* __restrict__ xnext, int A,
                                          int ld, int j0, int c0, int nvalid, double (&sxy)[V],
                                          double (&syy)[V], int colbase = 0) {
  constexpr int G = 4;
  const int kb = rowptr[A], ke = rowptr[A + 1];
  // A indexes rowptr (global block row); rows of the vectors are local: A - colbase
  const int Al = A - colbase;
  T* out = xnext + (size_t)(4 * Al) * ld + c0;
  T prev[4][V];
  if (MODE == 2) {
#pragma unroll
    for (int r = 0; r < 4; r++) load_row<T, V, NT>(out + (size_t)r * ld, prev[r]);
  }
  T acc[4][V], xa[4][V];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int v = 0; v < V; v++) acc[r][v] = T(0);
  for (int k0 = kb; k0 < ke; k0 += G) {
    T x[G][4][V];
#pragma unroll
    for (int u = 0; u < G; u++) {
      if (k0 + u < ke) {
        const int B = col[k0 + u] - colbase;
        if (MODE == 0) {
#pragma unroll
          for (int s = 0; s < 4; s++)
#pragma unroll
            for (int v = 0; v < V; v++)
              x[u][s][v] = (4 * B + s == j0 + c0 + v && c0 + v < nvalid) ? T(1) : T(0);
        } else {
#pragma unroll
          for (int s = 0; s < 4; s++)
            load_row<T, V, false>(xcur + (size_t)(4 * B + s) * ld + c0, x[u][s]);
        }
      }
    }
    if (k0 == kb) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int v = 0; v < V; v++) xa[s][v] = x[0][s][v];
    }
#pragma unroll
    for (int u = 0; u < G; u++) {
      if (k0 + u < ke) {
        const T* h = hv + 16 * (size_t)(k0 + u);
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
          for (int s = 0; s < 4; s++) {
            const T hrs = h[4 * r + s];
#pragma unroll
            for (int v = 0; v < V; v++) acc[r][v] = fma(hrs, x[u][s][v], acc[r][v]);
          }
      }
    }
  }
  T y[4][V];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int v = 0; v < V; v++) {
      if (MODE == 0 || MODE == 3)
        y[r][v] = acc[r][v];
      else if (MODE == 1)
        y[r][v] = T(2) * acc[r][v] - ((4 * Al + r == j0 + c0 + v && c0 + v < nvalid) ? T(1) : T(0));
      else
        y[r][v] = T(2) * acc[r][v] - prev[r][v];
    }
#pragma unroll
  for (int r = 0; r < 4; r++) store_row<T, V, NT>(out + (size_t)r * ld, y[r]);
  // the per-column dots are accumulated in FP64 (a product of two floats is
  // exact in double, so only the FP64 sums round)
#pragma unroll
  for (int v = 0; v < V; v++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const double yd = (double)y[r][v];
      sxy[v] = fma(yd, (double)xa[r][v], sxy[v]);
      syy[v] = fma(yd, yd, syy[v]);
    }
}

// sum the 4 waves' per-column partials in LDS (FP64) and write one row pair
template <int V>
__device__ __forceinline__ void flush_partials(const double (&sxy)[V], const double (&syy)[V], int wave,
                                               int lane, double* __restrict__ part, size_t row2,
                                               int ld, int colbase) {
  constexpr int W = 64 * V;
  __shared__ double sh[4][2][W];
#pragma unroll
  for (int v = 0; v < V; v++) {
    sh[wave][0][lane * V + v] = sxy[v];
    sh[wave][1][lane * V + v] = syy[v];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * W; e += 256) {
    const int which = e / W, cc = e - which * W;
    const double s = (sh[0][which][cc] + sh[1][which][cc]) + (sh[2][which][cc] + sh[3][which][cc]);
    part[(row2 + which) * ld + colbase + cc] = s;
  }
  __syncthreads();
}


}  // namespace ltbdev
