// Many independent structures of uneven size through one SCC: every
// structure keeps its own bounds, rescale, mu, populations, mixing and
// convergence; every recursion step of the exact-trace KPE is one launch for
// the whole batch.  No reference counterpart (the reference runs one structure
// per MPI job, testjob2.pbs); per structure the math is that of ltb_scc_run.
//
// Layout: atoms of all structures concatenated (BSR with global block
// columns, block-diagonal across structures).  Structure s keeps its whole
// exact trace in flight: T_n(s) is norb_s x ld_s (ld_s = norb_s rounded up to
// its column groups), row-major, at element offset xoff[s] of the ping-pong
// buffers of its current precision (FP32 and FP64 buffers are separate).
// A task is (s, column group, 4 apw atoms); a resident grid walks the task
// slots, skipping structures that are done or in the other precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "cheb_device.hpp"
#include "ctx.hpp"
#include "system.hpp"

namespace {

constexpr double kPi = 3.141592653589793238463;
enum { S_EMIN, S_EMAX, S_A, S_B, S_MU, S_NELEC, S_EBAND, S_EBFP, S_DQ, S_NE, S_BAD, NSC };

// Column groups of a structure: 256 wide (FP32: 4 columns per lane, 16-byte
// accesses; FP64: two passes of 2 per lane) and one tail group for the
// norb mod 256 columns left, 64, 128 or 256 wide (the narrowest that holds
// them; padding columns carry zeros).  ld_s = the columns of all groups.
__host__ __device__ inline int tail_width(int norb) {
  const int r = norb & 255;
  return r == 0 ? 0 : (r > 128 ? 256 : (r > 64 ? 128 : 64));
}

// One task = one column group of one structure x 4*apw consecutive atoms,
// described by one 64-byte record per grid-stride slot (s = -1: padding), so
// a workgroup fetches everything it needs about its next task with one
// scalar load, issued while it still works on the current one.
// Partials: 2 x 256 doubles per task.
// act: the structure's state copied in at the start of every SCC iteration
// (k_batch_mark): 0 done or padding, 1 FP32 recursion, 2 FP64 recursion; the
// FP32 and FP64 instances of the step kernel each take their own precision.
// Slot i runs on XCD i % 8 and every task of one column group shares an XCD.
struct alignas(64) BatchTask {
  long long xoff;   // element offset of the structure's vectors
  int s;            // structure, -1 for a padding slot
  int a0;           // its first atom (global block row)
  int nat;          // its atoms
  int ld;           // its row stride in elements (a power of two >= the columns of all its groups)
  int cstart;       // first column of the group
  int abase;        // first atom of the task, local to the structure
  int width;        // 256, 128 or 64 columns
  int row2;         // first of the task's two partial rows
  int act;          // active[s] of the current SCC iteration
  int pad[3];
};
static_assert(sizeof(BatchTask) == 64, "one scalar load per task");

__global__ __launch_bounds__(256) void k_batch_mark(BatchTask* __restrict__ tasks, int nslots,
                                                    const int* __restrict__ active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nslots) {
    const int s = tasks[i].s;
    tasks[i].act = s >= 0 ? active[s] : 0;
  }
}

// sum the 4 waves' per-column partials in LDS (FP64) and write one row pair;
// two LDS buffers used in turn, so one barrier per task suffices (a wave
// writing buffer p again has passed the barrier that every reader of its
// previous use reached after reading).  Task t owns 4 KB of partials: 2 x 256
// doubles (FP64 in both recursion precisions).
template <int V>
__device__ __forceinline__ void flush_task(const double (&sxy)[V], const double (&syy)[V], int wave, int lane,
                                           double (*sh)[2][256], double* __restrict__ part, size_t row2,
                                           int pcol) {
  constexpr int W = 64 * V;
#pragma unroll
  for (int v = 0; v < V; v++) {
    sh[wave][0][lane * V + v] = sxy[v];
    sh[wave][1][lane * V + v] = syy[v];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * W; e += 256) {
    const int which = e / W, cc = e - which * W;
    const double s = (sh[0][which][cc] + sh[1][which][cc]) + (sh[2][which][cc] + sh[3][which][cc]);
    part[(row2 + which) * 256 + pcol + cc] = s;
  }
}

template <typename T, int V, int MODE, bool NT>
__device__ __forceinline__ void task_pass(const int* __restrict__ rowptr, const int* __restrict__ col,
                                          const T* __restrict__ hv, const T* xc, T* xn, int a0, int ab,
                                          int ae, int ld, int c0, int nvalid, int wave, int lane,
                                          double (*sh)[2][256], double* __restrict__ part, size_t row2,
                                          int pcol) {
  double sxy[V], syy[V];
#pragma unroll
  for (int v = 0; v < V; v++) sxy[v] = syy[v] = 0.0;
  for (int A = ab; A < ae; A += 4)
    ltbdev::atom_step<T, V, MODE, NT>(rowptr, col, hv, xc, xn, a0 + A, ld, 0, c0, nvalid, sxy, syy, a0);
  flush_task<V>(sxy, syy, wave, lane, sh, part, row2, pcol);
}

template <typename T, int MODE, bool NT>
__global__ __launch_bounds__(256) void k_batch_step(const int* __restrict__ rowptr, const int* __restrict__ col,
                                                    const T* __restrict__ hv, const char* __restrict__ xcur,
                                                    char* __restrict__ xnext, int nslots,
                                                    const BatchTask* __restrict__ tasks,
                                                    double* __restrict__ part, int apw) {
  constexpr int want = sizeof(T) == 4 ? 1 : 2;
  __shared__ double sh[2][4][2][256];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  int i = blockIdx.x, buf = 0;
  if (i >= nslots) return;
  BatchTask cur = tasks[i];
  for (; i < nslots; i += gridDim.x) {
    const BatchTask nxt = tasks[min(i + (int)gridDim.x, nslots - 1)];  // in flight during this task
    if (cur.act == want) {
      const T* xc = reinterpret_cast<const T*>(xcur) + cur.xoff;
      T* xn = reinterpret_cast<T*>(xnext) + cur.xoff;
      const int nat = cur.nat, nvalid = 4 * nat;
      // the 4 waves take the task's atoms in turn, so a workgroup works on 4
      // consecutive atoms at a time
      const int ab = cur.abase + wave, ae = min(cur.abase + 4 * apw, nat);
      const size_t row2 = (size_t)cur.row2;
      if (cur.width == 256) {
        if constexpr (sizeof(T) == 4) {
          task_pass<T, 4, MODE, NT>(rowptr, col, hv, xc, xn, cur.a0, ab, ae, cur.ld, cur.cstart + 4 * lane,
                                    nvalid, wave, lane, sh[buf], part, row2, 0);
          buf ^= 1;
        } else {
          for (int h = 0; h < 2; h++) {
            task_pass<T, 2, MODE, NT>(rowptr, col, hv, xc, xn, cur.a0, ab, ae, cur.ld,
                                      cur.cstart + 128 * h + 2 * lane, nvalid, wave, lane, sh[buf], part, row2,
                                      128 * h);
            buf ^= 1;
          }
        }
      } else if (cur.width == 128) {
        task_pass<T, 2, MODE, NT>(rowptr, col, hv, xc, xn, cur.a0, ab, ae, cur.ld, cur.cstart + 2 * lane, nvalid,
                                  wave, lane, sh[buf], part, row2, 0);
        buf ^= 1;
      } else {
        task_pass<T, 1, MODE, NT>(rowptr, col, hv, xc, xn, cur.a0, ab, ae, cur.ld, cur.cstart + lane, nvalid,
                                  wave, lane, sh[buf], part, row2, 0);
        buf ^= 1;
      }
    }
    cur = nxt;
  }
}

// per column of every active structure: fold the task partials of step n into
// per-atom moments (as k_cheb_reduce for one structure)
__global__ __launch_bounds__(256) void k_batch_reduce(
    const double* __restrict__ part, int ncols, const int* __restrict__ gc_s,
    const int* __restrict__ gc_c, const int* __restrict__ tbegin, const int* __restrict__ nsg,
    const int* __restrict__ aoff, const int* __restrict__ active, int n, int M,
    double* __restrict__ mu1, double* __restrict__ mom, int nat_total) {
  const int gc = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int s = 0, c = 0;
  bool use = gc < ncols;
  if (use) {
    s = gc_s[gc];
    c = gc_c[gc];
    use = active[s] != 0;
  }
  double sxy = 0, syy = 0;
  bool valid = false;
  if (use) {
    const int nat = aoff[s + 1] - aoff[s];
    valid = c < 4 * nat;
    // 256-wide groups first, then the one tail group
    const int cw = (4 * nat) & ~255;
    const int gidx = c < cw ? c >> 8 : cw >> 8;
    const int off = c < cw ? c & 255 : c - cw;
    const int base = tbegin[s] + gidx * nsg[s];
    for (int g = 0; g < nsg[s]; g++) {
      sxy += part[(size_t)(base + g) * 512 + off];
      syy += part[(size_t)(base + g) * 512 + 256 + off];
    }
  }
  double modd = 0, meven = 0;
  if (use) {
    if (n == 0) {
      modd = sxy;
      mu1[gc] = sxy;
    } else {
      modd = 2.0 * sxy - mu1[gc];
    }
    meven = 2.0 * syy - (valid ? 1.0 : 0.0);
  }
  modd += __shfl_xor(modd, 1);
  modd += __shfl_xor(modd, 2);
  meven += __shfl_xor(meven, 1);
  meven += __shfl_xor(meven, 2);
  if (use && valid && (lane & 3) == 0) {
    const int atom = aoff[s] + (c >> 2);
    if (n == 0) mom[atom] = 4.0;
    if (2 * n + 1 <= M) mom[(size_t)(2 * n + 1) * nat_total + atom] = modd;
    if (2 * n + 2 <= M) mom[(size_t)(2 * n + 2) * nat_total + atom] = meven;
  }
}

__global__ __launch_bounds__(256) void k_batch_momtot(const double* __restrict__ mom, int nat_total,
                                                      const int* __restrict__ aoff,
                                                      const int* __restrict__ active, int M,
                                                      double* __restrict__ tot) {
  const int m = blockIdx.x, s = blockIdx.y;
  if (!active[s]) return;
  double acc = 0;
  for (int a = aoff[s] + threadIdx.x; a < aoff[s + 1]; a += 256) acc += mom[(size_t)m * nat_total + a];
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[(size_t)s * (M + 1) + m] = sh[0];
}

__global__ __launch_bounds__(256) void k_batch_gersh(const double* __restrict__ diag0,
                                                     const double* __restrict__ offabs,
                                                     const double* __restrict__ v,
                                                     const int* __restrict__ aoff,
                                                     const int* __restrict__ active,
                                                     double* __restrict__ sc) {
  const int s = blockIdx.x;
  if (!active[s]) return;
  double lo = INFINITY, hi = -INFINITY;
  int bad = 0;
  for (int r = 4 * aoff[s] + threadIdx.x; r < 4 * aoff[s + 1]; r += 256) {
    const double d = diag0[r] + v[r >> 2], o = offabs[r];
    if (!isfinite(d) || !isfinite(o)) bad = 1;
    lo = fmin(lo, d - o);
    hi = fmax(hi, d + o);
  }
  __shared__ double sl[256], sh[256];
  __shared__ int sb;
  if (threadIdx.x == 0) sb = 0;
  __syncthreads();
  if (bad) atomicOr(&sb, 1);
  sl[threadIdx.x] = lo;
  sh[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + w]);
      sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double* p = sc + (size_t)s * NSC;
    p[S_EMIN] = sl[0];
    p[S_EMAX] = sh[0];
    p[S_A] = (sh[0] - sl[0]) / 2;
    p[S_B] = (sh[0] + sl[0]) / 2;
    p[S_BAD] = (sb || !isfinite(sl[0]) || !isfinite(sh[0])) ? 1.0 : 0.0;
  }
}

__global__ void k_batch_rescale(const double* __restrict__ h0, const int* __restrict__ blkrow,
                                const int* __restrict__ col, const int* __restrict__ atomst,
                                const double* __restrict__ v, const double* __restrict__ sc,
                                const int* __restrict__ active, int nnzb, double* __restrict__ ht64,
                                float* __restrict__ ht32) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nnzb * 16) return;
  const int k = idx >> 4, e = idx & 15, row = blkrow[k], s = atomst[row];
  if (!active[s]) return;
  const double* p = sc + (size_t)s * NSC;
  double val = h0[idx];
  if (col[k] == row && (e % 5) == 0) val = (val + v[row]) - p[S_B];
  val = val / p[S_A];
  ht64[idx] = val;
  ht32[idx] = (float)val;
}

__device__ __forceinline__ double fermi(double x, double mu, double kT) {
  const double t = (x - mu) / kT;
  if (t > 0) {
    const double e = exp(-t);
    return e / (1.0 + e);
  }
  return 1.0 / (1.0 + exp(t));
}

__global__ void k_batch_weights(const double* __restrict__ tot, const double* __restrict__ g, int M,
                                int K, const int* __restrict__ active, double* __restrict__ W,
                                double* __restrict__ X) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, s = blockIdx.y;
  if (k >= K) return;
  const double th = kPi * (k + 0.5) / K;
  if (s == 0) X[k] = cos(th);
  if (!active[s]) return;
  const double* t = tot + (size_t)s * (M + 1);
  double acc = t[0];
  for (int m = 1; m < M; m++) acc += 2.0 * g[m] * t[m] * cos(m * th);
  W[(size_t)s * K + k] = acc / K;
}

__global__ __launch_bounds__(1024) void k_batch_bisect(const double* __restrict__ W,
                                                       const double* __restrict__ X, int K,
                                                       double kT, const int* __restrict__ active,
                                                       double* __restrict__ sc) {
  const int s = blockIdx.x;
  if (!active[s]) return;
  double* p = sc + (size_t)s * NSC;
  const double kTs = kT / p[S_A], Ne = p[S_NE];
  __shared__ double sh[1024];
  __shared__ double s_lo, s_hi;
  if (threadIdx.x == 0) {
    s_lo = -1.0;
    s_hi = 1.0;
  }
  __syncthreads();
  for (int it = 0; it < 200; it++) {
    const double lo = s_lo, hi = s_hi, mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    double part = 0;
    for (int k = threadIdx.x; k < K; k += 1024) part += fermi(X[k], mid, kTs) * W[(size_t)s * K + k];
    sh[threadIdx.x] = part;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (2.0 * sh[0] < Ne)
        s_lo = mid;
      else
        s_hi = mid;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) p[S_MU] = 0.5 * (s_lo + s_hi);
}

__global__ __launch_bounds__(256) void k_batch_coeffs(const double* __restrict__ X, int K,
                                                      const double* __restrict__ g, int M, double kT,
                                                      const int* __restrict__ active,
                                                      const double* __restrict__ sc,
                                                      double* __restrict__ w) {
  const int m = blockIdx.x, s = blockIdx.y;
  if (!active[s]) return;
  const double* p = sc + (size_t)s * NSC;
  const double mu = p[S_MU], kTs = kT / p[S_A];
  double acc = 0;
  for (int k = threadIdx.x; k < K; k += 256) {
    const double th = kPi * (k + 0.5) / K;
    acc += fermi(X[k], mu, kTs) * cos(m * th);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) sh[threadIdx.x] += sh[threadIdx.x + q];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double c = sh[0] * (2.0 / K);
    w[(size_t)s * M + m] = m == 0 ? c / 2.0 : c * g[m];
  }
}

__global__ void k_batch_pops(const double* __restrict__ mom, const double* __restrict__ w, int M,
                             int nat_total, const int* __restrict__ atomst,
                             const int* __restrict__ active, double* __restrict__ pop) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat_total) return;
  const int s = atomst[a];
  if (!active[s]) return;
  double acc = 0;
  for (int m = 0; m < M; m++) acc += w[(size_t)s * M + m] * mom[(size_t)m * nat_total + a];
  pop[a] = acc;
}

// per structure: electrons, band energy, |dq|, first-order fixed-point band
// energy, then mixing of the structure's populations
__global__ __launch_bounds__(256) void k_batch_finish(
    const double* __restrict__ tot, const double* __restrict__ w, int M, const int* __restrict__ aoff,
    const int* __restrict__ grow, const int* __restrict__ gcol, const double* __restrict__ gval,
    double* __restrict__ pin, const double* __restrict__ pout, double mix,
    const int* __restrict__ active, double* __restrict__ sc) {
  const int s = blockIdx.x;
  if (!active[s]) return;
  double* p = sc + (size_t)s * NSC;
  double mx = 0, corr = 0;
  int bad = 0;
  for (int a = aoff[s] + threadIdx.x; a < aoff[s + 1]; a += 256) {
    const double d = 2.0 * fabs(pout[a] - pin[a]);
    if (!isfinite(d)) bad = 1;
    mx = fmax(mx, d);
    double gd = 0;
    for (int k = grow[a]; k < grow[a + 1]; k++) gd += gval[k] * (pout[gcol[k]] - pin[gcol[k]]);
    corr += pout[a] * gd;
  }
  __shared__ double smx[256], scr[256];
  __shared__ int sb;
  if (threadIdx.x == 0) sb = 0;
  __syncthreads();
  if (bad) atomicOr(&sb, 1);
  smx[threadIdx.x] = mx;
  scr[threadIdx.x] = corr;
  __syncthreads();
  for (int q = 128; q > 0; q >>= 1) {
    if ((int)threadIdx.x < q) {
      smx[threadIdx.x] = fmax(smx[threadIdx.x], smx[threadIdx.x + q]);
      scr[threadIdx.x] += scr[threadIdx.x + q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double* t = tot + (size_t)s * (M + 1);
    const double* ws = w + (size_t)s * M;
    double ns = 0, tph = ws[0] * t[1];
    for (int m = 0; m < M; m++) ns += ws[m] * t[m];
    for (int m = 1; m < M; m++) tph += ws[m] * 0.5 * (t[m + 1] + t[m - 1]);
    p[S_NELEC] = 2.0 * ns;
    p[S_EBAND] = 2.0 * (p[S_A] * tph + p[S_B] * ns);
    p[S_EBFP] = p[S_EBAND] + 2.0 * scr[0];
    p[S_DQ] = smx[0];
    if (sb || !isfinite(p[S_EBAND])) p[S_BAD] = 1.0;
  }
  __syncthreads();
  for (int a = aoff[s] + threadIdx.x; a < aoff[s + 1]; a += 256) pin[a] += mix * (pout[a] - pin[a]);
}

__global__ void k_batch_shift(const int* __restrict__ grow, const int* __restrict__ gcol,
                              const double* __restrict__ gval, const double* __restrict__ p,
                              const double* __restrict__ p0, const int* __restrict__ atomst,
                              const int* __restrict__ active, double* __restrict__ v, int nat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nat || !active[atomst[i]]) return;
  double acc = 0;
  for (int k = grow[i]; k < grow[i + 1]; k++) acc += gval[k] * (p[gcol[k]] - p0[gcol[k]]);
  v[i] = acc;
}

// two-level SCC, per atom by its structure's mode (1 low-order iteration,
// 2 full-order iteration): after the order-M_lo populations, a low-order
// iteration adds the defect and a full-order one keeps them
__global__ void k_batch_lo_combine(double* __restrict__ pout, double* __restrict__ plo,
                                   const double* __restrict__ defect, const int* __restrict__ atomst,
                                   const int* __restrict__ mode, int nat) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nat) return;
  const int md = mode[atomst[a]];
  if (md == 1) pout[a] += defect[a];
  if (md == 2) plo[a] = pout[a];
}

// after the order-M populations: the defect of a full-order iteration
__global__ void k_batch_defect(const double* __restrict__ pout, const double* __restrict__ plo,
                               double* __restrict__ defect, const int* __restrict__ atomst,
                               const int* __restrict__ mode, int nat) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < nat && mode[atomst[a]] == 2) defect[a] = pout[a] - plo[a];
}

void kernel_gmM(int M, std::vector<double>& g) {
  std::vector<double> U(M);
  const double alpha = kPi / (M + 2), lam = std::cos(alpha);
  double nrm = 0;
  for (int m = 0; m < M; m++) {
    U[m] = std::sin((m + 1) * std::acos(lam)) / std::sin(std::acos(lam));
    nrm += U[m] * U[m];
  }
  nrm = std::sqrt(nrm);
  for (double& u : U) u /= nrm;
  g.assign(M, 0.0);
  for (int m = 0; m < M; m++)
    for (int i = 0; i < M - m - 1; i++) g[m] += U[i] * U[m + i];
}

}  // namespace

struct ltb_batch {
  ltb_ctx* ctx = nullptr;
  int S = 0, nat = 0, nnzb = 0, ntasks = 0, ncols = 0;
  std::vector<int> aoff, lds;  // per structure: atom offsets, columns of all its groups
  std::vector<int> strides;    // per structure: row stride of its vectors (lds rounded up to a power of two)
  std::vector<size_t> xoff;
  std::vector<void*> allocs;
  int *d_rowptr, *d_col, *d_blkrow, *d_atomst, *d_grow, *d_gcol, *d_aoff, *d_active;
  int nslots = 0;
  ltb_kpe_tiling tiling{256, 4, LTB_SCHED_SWEEP, 8, 8, 1, 2, 0};  // column group, V, -, blocks/CU, atoms/wave, NT, index order, -
  void* d_tasks = nullptr;  // BatchTask per slot
  int *d_tbegin = nullptr, *d_nsg = nullptr;
  int *d_gc_s, *d_gc_c;
  double *d_h0, *d_ht64, *d_diag0, *d_offabs, *d_gval, *d_p0, *d_pin, *d_pout, *d_v, *d_sc;
  float* d_ht32;
  size_t xelems = 0;
  // ping-pong vector buffers per precision; FP32 ones packed on their own
  // (half the address range of FP64 room, which measurably helps the step)
  void* d_x32[2] = {nullptr, nullptr};
  void* d_x64[2] = {nullptr, nullptr};
  double *d_part = nullptr, *d_mu1 = nullptr, *d_mom = nullptr, *d_tot = nullptr, *d_W = nullptr,
         *d_X = nullptr, *d_g = nullptr, *d_w = nullptr;
  int M = -1, K = -1;
  // two-level SCC: the order-M_lo quadrature state, per-structure masks
  // (act_full: recursing past step M_lo/2 this iteration; act_low: finishing
  // at order M_lo; mode: 1 low-order, 2 full-order two-level iteration) and
  // per-atom low-order populations and defect
  double *d_tot_lo = nullptr, *d_g_lo = nullptr, *d_w_lo = nullptr;
  int M_lo = -1;
  int* d_masks = nullptr;  // [3][S]: act_full, act_low, mode
  double *d_plo = nullptr, *d_defect = nullptr;
  std::vector<double> Ne;
  std::vector<int> nnzb_s;
  // step-kernel statistics of the last ltb_batch_scc_run (events only in eager
  // iterations while the context's KPE timing is on)
  std::vector<hipEvent_t> events;
  long long stat_launches = 0;
  double stat_seconds = 0, stat_bytes = 0;
};

namespace {

template <typename P>
int bmalloc(ltb_batch* b, P** p, size_t bytes) {
  hipError_t e = hipMalloc((void**)p, std::max<size_t>(bytes, 8));
  if (e != hipSuccess) return set_err(b->ctx, LTB_ERR_NOMEM, "batch allocation failed");
  b->allocs.push_back((void*)*p);
  return LTB_OK;
}

template <typename P>
int bupload(ltb_batch* b, P** p, const std::vector<P>& h) {
  int st = bmalloc(b, p, sizeof(P) * h.size());
  if (st) return st;
  if (!h.empty() && hipMemcpy(*p, h.data(), sizeof(P) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_DEVICE, "batch upload failed");
  return LTB_OK;
}

// algorithmic bytes of the step kernels of one iteration: per structure and
// step, the vectors read by the gathers (t_n), read (t_{n-1}) and written
// (t_{n+1}) once over the valid columns (1, 2, 3 streams for steps 0, 1, >1),
// plus H~ blocks, block-column indices and row pointers once
double iteration_bytes(const ltb_batch* b, const std::vector<int>& active, int n0, int n1) {
  double bytes = 0;
  for (int s = 0; s < b->S; s++) {
    if (!active[s]) continue;
    const double el = active[s] == 1 ? 4.0 : 8.0;
    const double norb = 4.0 * (b->aoff[s + 1] - b->aoff[s]);
    const double hb = b->nnzb_s[s] * (16.0 * el + 4.0) + 4.0 * (norb / 4 + 1);
    for (int n = n0; n < n1; n++) {
      const int streams = n == 0 ? 1 : (n == 1 ? 2 : 3);
      bytes += streams * el * norb * norb + hb;
    }
  }
  return bytes;
}

void free_tasks(ltb_batch* b) {
  for (int** p : {&b->d_tbegin, &b->d_nsg}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  if (b->d_tasks) (void)hipFree(b->d_tasks);
  b->d_tasks = nullptr;
  if (b->d_part) (void)hipFree(b->d_part);
  b->d_part = nullptr;
}

int upload_tmp(ltb_batch* b, int** p, const std::vector<int>& h) {
  if (hipMalloc((void**)p, std::max<size_t>(sizeof(int) * h.size(), 8)) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_NOMEM, "batch task allocation failed");
  if (!h.empty() && hipMemcpy(*p, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_DEVICE, "batch task upload failed");
  return LTB_OK;
}

// Tasks for the batch's tiling: per structure its column groups (256-wide,
// then the tail group), each split into blocks of
// 4*atoms_per_wave atoms.  Column groups go to the least loaded XCD (load =
// columns x atoms) so that each XCD's L2 serves the gathers of its own
// groups; slot 8 i + x holds XCD x's i-th task.  Task ids (the partial row
// pairs) run over structures, then column groups, then atom blocks, which is
// the order k_batch_reduce reads them in.
int build_tasks(ltb_batch* b) {
  const int apw = b->tiling.atoms_per_wave;
  std::vector<int> tbegin, nsg;
  std::vector<BatchTask> all;
  std::vector<std::vector<int>> xcd_tasks(8);
  std::vector<double> xcd_load(8, 0.0);
  for (int s = 0; s < b->S; s++) {
    const int nat = b->aoff[s + 1] - b->aoff[s], norb = 4 * nat, ld = b->lds[s];
    const int ng = (nat + 4 * apw - 1) / (4 * apw), cw = norb & ~255, tw = tail_width(norb);
    tbegin.push_back((int)all.size());
    nsg.push_back(ng);
    for (int c0 = 0; c0 < ld;) {
      const int width = c0 < cw ? 256 : tw;
      const int x = (int)(std::min_element(xcd_load.begin(), xcd_load.end()) - xcd_load.begin());
      xcd_load[x] += (double)width * nat;
      for (int g = 0; g < ng; g++) {
        BatchTask t{};
        t.xoff = (long long)b->xoff[s];
        t.s = s;
        t.a0 = b->aoff[s];
        t.nat = nat;
        t.ld = b->strides[s];
        t.cstart = c0;
        t.abase = g * 4 * apw;
        t.width = width;
        t.row2 = 2 * (int)all.size();
        xcd_tasks[x].push_back((int)all.size());
        all.push_back(t);
      }
      c0 += width;
    }
  }
  size_t per_xcd = 0;
  for (auto& v : xcd_tasks) per_xcd = std::max(per_xcd, v.size());
  BatchTask pad{};
  pad.s = -1;
  std::vector<BatchTask> slots(8 * per_xcd, pad);
  for (int x = 0; x < 8; x++)
    for (size_t i = 0; i < xcd_tasks[x].size(); i++) slots[8 * i + x] = all[xcd_tasks[x][i]];
  free_tasks(b);
  b->nslots = (int)slots.size();
  b->ntasks = (int)all.size();
  int st;
  if ((st = upload_tmp(b, &b->d_tbegin, tbegin)) || (st = upload_tmp(b, &b->d_nsg, nsg))) return st;
  if (hipMalloc(&b->d_tasks, sizeof(BatchTask) * std::max<size_t>(slots.size(), 1)) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_NOMEM, "batch task allocation failed");
  if (!slots.empty() &&
      hipMemcpy(b->d_tasks, slots.data(), sizeof(BatchTask) * slots.size(), hipMemcpyHostToDevice) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_DEVICE, "batch task upload failed");
  // the structures' current states (every SCC iteration copies them in again)
  hipLaunchKernelGGL(k_batch_mark, dim3((b->nslots + 255) / 256), dim3(256), 0, b->ctx->stream,
                     (BatchTask*)b->d_tasks, b->nslots, b->d_active);
  if (hipGetLastError() != hipSuccess) return set_err(b->ctx, LTB_ERR_DEVICE, "batch task marking failed");
  if (hipMalloc((void**)&b->d_part, sizeof(double) * 512 * (size_t)b->ntasks) != hipSuccess)
    return set_err(b->ctx, LTB_ERR_NOMEM, "batch partials allocation failed");
  return LTB_OK;
}

// ping-pong vector buffers of both precisions for every structure (a
// structure recurses in FP32 or FP64 from one SCC iteration to the next)
int ensure_vectors(ltb_batch* b) {
  ltb_ctx* c = b->ctx;
  for (int k = 0; k < 2; k++) {
    if (!b->d_x32[k]) LTB_HIP(c, hipMalloc(&b->d_x32[k], std::max<size_t>(b->xelems * sizeof(float), 8)));
    if (!b->d_x64[k]) LTB_HIP(c, hipMalloc(&b->d_x64[k], std::max<size_t>(b->xelems * sizeof(double), 8)));
  }
  return LTB_OK;
}

// buffers of recursion step n: t_n is read from cur, t_{n+1} written over
// t_{n-1} in nxt
template <typename T>
void step_buffers(const ltb_batch* b, int n, void** cur, void** nxt) {
  void* const* x = sizeof(T) == 4 ? b->d_x32 : b->d_x64;
  *cur = x[n & 1];
  *nxt = n == 0 ? x[1] : x[(n & 1) ^ 1];
}

// resident grid: a multiple of 8 blocks keeps slot i on XCD i % 8 through the
// grid-stride loop
int step_grid(const ltb_batch* b) {
  const int per_cu = std::max(1, b->tiling.blocks_per_cu);
  const int grid = std::min(b->nslots, per_cu * std::max(b->ctx->cu_count, 8));
  return std::max(8, grid / 8 * 8);
}

template <typename T, int MODE>
void launch_step(ltb_batch* b, int grid, const void* cur, void* nxt) {
  const T* hv = sizeof(T) == 4 ? (const T*)b->d_ht32 : (const T*)b->d_ht64;
  auto kern = b->tiling.nontemporal ? k_batch_step<T, MODE, true> : k_batch_step<T, MODE, false>;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, b->ctx->stream, b->d_rowptr, b->d_col, hv,
                     (const char*)cur, (char*)nxt, b->nslots, (const BatchTask*)b->d_tasks, b->d_part,
                     b->tiling.atoms_per_wave);
}

template <typename T>
void launch_step(ltb_batch* b, int mode, int grid, const void* cur, void* nxt) {
  if (mode == 0)
    launch_step<T, 0>(b, grid, cur, nxt);
  else if (mode == 1)
    launch_step<T, 1>(b, grid, cur, nxt);
  else
    launch_step<T, 2>(b, grid, cur, nxt);
}

// recursion steps n0 <= n < n1 of the exact-trace moments (order M) of the
// structures active in `act` (the tasks are marked from it first); `precs`
// bit 0: some structure recurses in FP32, bit 1: some in FP64
int batch_steps(ltb_batch* b, int M, int n0, int n1, int precs, const int* act, bool capturing) {
  ltb_ctx* c = b->ctx;
  const int nsteps = (M + 1) / 2;
  const int grid = step_grid(b);
  const bool timed = c->timing && !capturing;
  if (timed) {
    while ((int)b->events.size() < 2 * nsteps) {
      hipEvent_t e;
      LTB_HIP(c, hipEventCreate(&e));
      b->events.push_back(e);
    }
  }
  hipLaunchKernelGGL(k_batch_mark, dim3((b->nslots + 255) / 256), dim3(256), 0, c->stream,
                     (BatchTask*)b->d_tasks, b->nslots, act);
  for (int n = n0; n < n1; n++) {
    const int mode = n == 0 ? 0 : (n == 1 ? 1 : 2);
    void *cur, *nxt;
    if (timed) LTB_HIP(c, hipEventRecord(b->events[2 * n], c->stream));
    if (precs & 1) {
      step_buffers<float>(b, n, &cur, &nxt);
      launch_step<float>(b, mode, grid, cur, nxt);
    }
    if (precs & 2) {
      step_buffers<double>(b, n, &cur, &nxt);
      launch_step<double>(b, mode, grid, cur, nxt);
    }
    if (timed) LTB_HIP(c, hipEventRecord(b->events[2 * n + 1], c->stream));
    hipLaunchKernelGGL(k_batch_reduce, dim3((b->ncols + 255) / 256), dim3(256), 0, c->stream, b->d_part,
                       b->ncols, b->d_gc_s, b->d_gc_c, b->d_tbegin, b->d_nsg, b->d_aoff, act, n,
                       M, b->d_mu1, b->d_mom, b->nat);
  }
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

// from the moment totals to the populations at order M of the structures
// active in `act`, then their |dq|, band energy and mixing
void batch_density(ltb_batch* b, const ltb_scc_params* prm, int M, const double* tot, const double* g, double* w,
                   const int* act) {
  ltb_ctx* c = b->ctx;
  const int K = prm->K;
  hipLaunchKernelGGL(k_batch_weights, dim3((K + 255) / 256, b->S), dim3(256), 0, c->stream, tot, g, M, K, act,
                     b->d_W, b->d_X);
  hipLaunchKernelGGL(k_batch_bisect, dim3(b->S), dim3(1024), 0, c->stream, b->d_W, b->d_X, K, prm->kT, act,
                     b->d_sc);
  hipLaunchKernelGGL(k_batch_coeffs, dim3(M, b->S), dim3(256), 0, c->stream, b->d_X, K, g, M, prm->kT, act,
                     b->d_sc, w);
  hipLaunchKernelGGL(k_batch_pops, dim3((b->nat + 255) / 256), dim3(256), 0, c->stream, b->d_mom, w, M, b->nat,
                     b->d_atomst, act, b->d_pout);
}

void batch_finish(ltb_batch* b, const ltb_scc_params* prm, int M, const double* tot, const double* w,
                  const int* act) {
  hipLaunchKernelGGL(k_batch_finish, dim3(b->S), dim3(256), 0, b->ctx->stream, tot, w, M, b->d_aoff, b->d_grow,
                     b->d_gcol, b->d_gval, b->d_pin, b->d_pout, prm->mix, act, b->d_sc);
}

// One SCC iteration of the batch.  precs / precs_full: precisions of the
// structures recursing up to step M_lo/2 / past it (two-level only; without
// it every structure recurses to M and precs_full is unused).
int batch_iteration(ltb_batch* b, const ltb_scc_params* prm, int precs, int precs_full, bool capturing) {
  ltb_ctx* c = b->ctx;
  const int M = prm->M, nsteps = (M + 1) / 2;
  hipLaunchKernelGGL(k_batch_shift, dim3((b->nat + 255) / 256), dim3(256), 0, c->stream, b->d_grow,
                     b->d_gcol, b->d_gval, b->d_pin, b->d_p0, b->d_atomst, b->d_active, b->d_v, b->nat);
  hipLaunchKernelGGL(k_batch_gersh, dim3(b->S), dim3(256), 0, c->stream, b->d_diag0, b->d_offabs, b->d_v,
                     b->d_aoff, b->d_active, b->d_sc);
  hipLaunchKernelGGL(k_batch_rescale, dim3((b->nnzb * 16 + 255) / 256), dim3(256), 0, c->stream, b->d_h0,
                     b->d_blkrow, b->d_col, b->d_atomst, b->d_v, b->d_sc, b->d_active, b->nnzb, b->d_ht64,
                     b->d_ht32);
  int st;
  if (prm->M_lo <= 0) {
    if ((st = batch_steps(b, M, 0, nsteps, precs, b->d_active, capturing))) return st;
    hipLaunchKernelGGL(k_batch_momtot, dim3(M + 1, b->S), dim3(256), 0, c->stream, b->d_mom, b->nat,
                       b->d_aoff, b->d_active, M, b->d_tot);
    batch_density(b, prm, M, b->d_tot, b->d_g, b->d_w, b->d_active);
    batch_finish(b, prm, M, b->d_tot, b->d_w, b->d_active);
    LTB_HIP(c, hipGetLastError());
    return LTB_OK;
  }
  // two-level: every structure to step M_lo/2, those in a full-order (or
  // plain) iteration on to M/2; order-M_lo populations for all (a low-order
  // iteration adds the defect and finishes there), order M for the others
  const int M_lo = prm->M_lo, nlo = (M_lo + 1) / 2;
  const int *act_full = b->d_masks, *act_low = b->d_masks + b->S, *mode = b->d_masks + 2 * b->S;
  const dim3 ga((b->nat + 255) / 256);
  if ((st = batch_steps(b, M, 0, nlo, precs, b->d_active, capturing))) return st;
  if (precs_full && (st = batch_steps(b, M, nlo, nsteps, precs_full, act_full, capturing))) return st;
  hipLaunchKernelGGL(k_batch_momtot, dim3(M_lo + 1, b->S), dim3(256), 0, c->stream, b->d_mom, b->nat,
                     b->d_aoff, b->d_active, M_lo, b->d_tot_lo);
  batch_density(b, prm, M_lo, b->d_tot_lo, b->d_g_lo, b->d_w_lo, b->d_active);
  hipLaunchKernelGGL(k_batch_lo_combine, ga, dim3(256), 0, c->stream, b->d_pout, b->d_plo, b->d_defect,
                     b->d_atomst, mode, b->nat);
  batch_finish(b, prm, M_lo, b->d_tot_lo, b->d_w_lo, act_low);
  if (precs_full) {
    hipLaunchKernelGGL(k_batch_momtot, dim3(M + 1, b->S), dim3(256), 0, c->stream, b->d_mom, b->nat,
                       b->d_aoff, act_full, M, b->d_tot);
    batch_density(b, prm, M, b->d_tot, b->d_g, b->d_w, act_full);
    hipLaunchKernelGGL(k_batch_defect, ga, dim3(256), 0, c->stream, b->d_pout, b->d_plo, b->d_defect,
                       b->d_atomst, mode, b->nat);
    batch_finish(b, prm, M, b->d_tot, b->d_w, act_full);
  }
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

}  // namespace

extern "C" {

int ltb_batch_create(ltb_ctx* c, int nstruct, const ltb_system* const* systems, ltb_batch** out) {
  if (!c || nstruct < 1 || !systems || !out) return LTB_ERR_ARG;
  *out = nullptr;
  auto* b = new ltb_batch();
  b->ctx = c;
  b->S = nstruct;
  std::vector<int> rowptr{0}, col, blkrow, atomst, grow{0}, gcol;
  std::vector<double> h0, diag0, offabs, gval, p0;
  b->aoff.push_back(0);
  for (int s = 0; s < nstruct; s++) {
    const ltb_system* sy = systems[s];
    if (!sy) return delete b, set_err(c, LTB_ERR_ARG, "null system in batch");
    const int a0 = b->aoff.back(), n = sy->natoms;
    for (int i = 0; i < n; i++) {
      for (int k = sy->h_rowptr[i]; k < sy->h_rowptr[i + 1]; k++) {
        col.push_back(a0 + sy->h_col[k]);
        blkrow.push_back(a0 + i);
        h0.insert(h0.end(), sy->h_val.begin() + 16 * (size_t)k, sy->h_val.begin() + 16 * (size_t)k + 16);
      }
      rowptr.push_back((int)col.size());
      for (int k = sy->g_rowptr[i]; k < sy->g_rowptr[i + 1]; k++) {
        gcol.push_back(a0 + sy->g_col[k]);
        gval.push_back(sy->g_val[k]);
      }
      grow.push_back((int)gcol.size());
      atomst.push_back(s);
      p0.push_back(sy->p0[i]);
      for (int a = 0; a < 4; a++) {
        double dg = 0, off = 0;
        for (int k = sy->h_rowptr[i]; k < sy->h_rowptr[i + 1]; k++)
          for (int q = 0; q < 4; q++) {
            const double v = sy->h_val[16 * (size_t)k + 4 * a + q];
            if (sy->h_col[k] == i && q == a)
              dg = v;
            else
              off += std::fabs(v);
          }
        diag0.push_back(dg);
        offabs.push_back(off);
      }
    }
    b->aoff.push_back(a0 + n);
    b->Ne.push_back(sy->nelectrons);
    b->nnzb_s.push_back(sy->h_rowptr[n]);
  }
  b->nat = b->aoff.back();
  b->nnzb = (int)col.size();
  // vectors and columns (tasks depend on the tiling: build_tasks)
  std::vector<int> gc_s, gc_c;
  b->xoff.push_back(0);
  for (int s = 0; s < nstruct; s++) {
    const int nat = b->aoff[s + 1] - b->aoff[s], norb = 4 * nat, ld = (norb & ~255) + tail_width(norb);
    int stride = 256;
    while (stride < ld) stride *= 2;
    b->lds.push_back(ld);
    b->strides.push_back(stride);
    b->xoff.push_back(b->xoff.back() + (size_t)norb * stride);
    for (int cc = 0; cc < ld; cc++) {
      gc_s.push_back(s);
      gc_c.push_back(cc);
    }
  }
  b->xelems = b->xoff.back();
  b->ncols = (int)gc_s.size();
  std::vector<int> ones(nstruct, 1);
  std::vector<double> sc((size_t)nstruct * NSC, 0.0);
  for (int s = 0; s < nstruct; s++) sc[(size_t)s * NSC + S_NE] = b->Ne[s];
  int st;
  if ((st = bupload(b, &b->d_rowptr, rowptr)) || (st = bupload(b, &b->d_col, col)) ||
      (st = bupload(b, &b->d_blkrow, blkrow)) || (st = bupload(b, &b->d_atomst, atomst)) ||
      (st = bupload(b, &b->d_grow, grow)) || (st = bupload(b, &b->d_gcol, gcol)) ||
      (st = bupload(b, &b->d_aoff, b->aoff)) ||
      (st = bupload(b, &b->d_active, ones)) || (st = bupload(b, &b->d_gc_s, gc_s)) || (st = bupload(b, &b->d_gc_c, gc_c)) ||
      (st = bupload(b, &b->d_h0, h0)) ||
      (st = bupload(b, &b->d_diag0, diag0)) || (st = bupload(b, &b->d_offabs, offabs)) ||
      (st = bupload(b, &b->d_gval, gval)) || (st = bupload(b, &b->d_p0, p0)) ||
      (st = bupload(b, &b->d_pin, p0)) || (st = bupload(b, &b->d_pout, p0)) ||
      (st = bupload(b, &b->d_sc, sc)) || (st = bmalloc(b, &b->d_v, sizeof(double) * b->nat)) ||
      (st = bmalloc(b, &b->d_ht64, sizeof(double) * 16 * b->nnzb)) ||
      (st = bmalloc(b, &b->d_ht32, sizeof(float) * 16 * b->nnzb)) ||
      (st = bmalloc(b, &b->d_mu1, sizeof(double) * b->ncols)) || (st = build_tasks(b))) {
    ltb_batch_destroy(b);
    return st;
  }
  *out = b;
  return LTB_OK;
}

int ltb_batch_sizes(const ltb_batch* b, int* nstruct, int* natoms_total, int* atom_offsets) {
  if (!b) return LTB_ERR_ARG;
  if (nstruct) *nstruct = b->S;
  if (natoms_total) *natoms_total = b->nat;
  if (atom_offsets) std::copy(b->aoff.begin(), b->aoff.end(), atom_offsets);
  return LTB_OK;
}

int ltb_batch_stats(const ltb_batch* b, long long* launches, double* seconds, double* bytes) {
  if (!b) return LTB_ERR_ARG;
  if (launches) *launches = b->stat_launches;
  if (seconds) *seconds = b->stat_seconds;
  if (bytes) *bytes = b->stat_bytes;
  return LTB_OK;
}

void ltb_batch_destroy(ltb_batch* b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->ctx->stream);
  for (void* p : b->allocs) (void)hipFree(p);
  free_tasks(b);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (void* p : {b->d_x32[0], b->d_x32[1], b->d_x64[0], b->d_x64[1], (void*)b->d_mom, (void*)b->d_tot, (void*)b->d_W,
                  (void*)b->d_X, (void*)b->d_g, (void*)b->d_w, (void*)b->d_tot_lo, (void*)b->d_g_lo,
                  (void*)b->d_w_lo})
    if (p) (void)hipFree(p);
  delete b;
}

int ltb_batch_scc_run(ltb_batch* b, const ltb_scc_params* prm, const double* guess, double* p_out,
                      int* iters, double* eband_fp, double* mu_ha, int* status, ltb_scc_result* total) {
  if (!b || !prm) return LTB_ERR_ARG;
  ltb_ctx* c = b->ctx;
  if (prm->M < 2 || !(prm->kT > 0) || prm->K < 16 || prm->max_iter < 1 || !(prm->tol > 0) ||
      !(prm->mix > 0) || prm->mix > 1)
    return set_err(c, LTB_ERR_ARG, "bad batch SCC parameters (Fermi-Dirac kT > 0 required)");
  if (prm->M_lo < 0 || (prm->M_lo > 0 && (prm->M_lo < 2 || prm->M_lo >= prm->M)) || !(prm->lo_tol >= 0))
    return set_err(c, LTB_ERR_ARG, "bad two-level parameters (0 or 2 <= M_lo < M, lo_tol >= 0)");
  const int M = prm->M, K = prm->K, S = b->S;
  const bool two = prm->M_lo > 0;
  auto t0 = std::chrono::steady_clock::now();
  int st0 = ensure_vectors(b);
  if (st0) return st0;
  if (b->M != M || b->K != K) {
    for (void** p : {(void**)&b->d_mom, (void**)&b->d_tot, (void**)&b->d_W, (void**)&b->d_X, (void**)&b->d_g,
                     (void**)&b->d_w}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    b->M = b->K = 0;  // no set looks valid until every allocation below succeeded
    LTB_HIP(c, hipMalloc(&b->d_mom, sizeof(double) * (size_t)(M + 1) * b->nat));
    LTB_HIP(c, hipMalloc(&b->d_tot, sizeof(double) * (size_t)(M + 1) * S));
    LTB_HIP(c, hipMalloc(&b->d_W, sizeof(double) * (size_t)K * S));
    LTB_HIP(c, hipMalloc(&b->d_X, sizeof(double) * K));
    LTB_HIP(c, hipMalloc(&b->d_g, sizeof(double) * M));
    LTB_HIP(c, hipMalloc(&b->d_w, sizeof(double) * (size_t)M * S));
    std::vector<double> g;
    kernel_gmM(M, g);
    LTB_HIP(c, hipMemcpy(b->d_g, g.data(), sizeof(double) * M, hipMemcpyHostToDevice));
    b->M = M;
    b->K = K;
  }
  if (two && b->M_lo != prm->M_lo) {
    const int L = prm->M_lo;
    for (void** p : {(void**)&b->d_tot_lo, (void**)&b->d_g_lo, (void**)&b->d_w_lo}) {
      if (*p) (void)hipFree(*p);
      *p = nullptr;
    }
    b->M_lo = 0;
    LTB_HIP(c, hipMalloc(&b->d_tot_lo, sizeof(double) * (size_t)(L + 1) * S));
    LTB_HIP(c, hipMalloc(&b->d_g_lo, sizeof(double) * L));
    LTB_HIP(c, hipMalloc(&b->d_w_lo, sizeof(double) * (size_t)L * S));
    std::vector<double> g;
    kernel_gmM(L, g);
    LTB_HIP(c, hipMemcpy(b->d_g_lo, g.data(), sizeof(double) * L, hipMemcpyHostToDevice));
    b->M_lo = L;
  }
  if (two && (!b->d_masks || !b->d_plo || !b->d_defect)) {
    int st1;
    if ((!b->d_masks && (st1 = bmalloc(b, &b->d_masks, sizeof(int) * 3 * (size_t)S))) ||
        (!b->d_plo && (st1 = bmalloc(b, &b->d_plo, sizeof(double) * b->nat))) ||
        (!b->d_defect && (st1 = bmalloc(b, &b->d_defect, sizeof(double) * b->nat))))
      return st1;
  }
  if (guess)
    LTB_HIP(c, hipMemcpy(b->d_pin, guess, sizeof(double) * b->nat, hipMemcpyHostToDevice));
  else
    LTB_HIP(c, hipMemcpy(b->d_pin, b->d_p0, sizeof(double) * b->nat, hipMemcpyDeviceToDevice));
  // per structure: 1 recursing in FP32, 2 in FP64, 0 done (converged or failed)
  const int start = prm->precision == LTB_FP64 ? 2 : 1;
  std::vector<int> active(S, start), it_s(S, 0), st_s(S, LTB_NOT_CONVERGED);
  // two-level, per structure in FP32: the phase (true: low-order iterations)
  // and its stall test (as ltb_scc_run); a structure whose full-order |dq|
  // falls below switch_tol goes on in plain FP64 iterations
  std::vector<char> low(S, 0);
  std::vector<int> plen(S, 0), stall(S, 0);
  std::vector<double> best(S, 0.0), dq_full(S, 0.0);
  const double lo_tol = prm->lo_tol > 0 ? prm->lo_tol : 0.1 * prm->tol;
  std::vector<int> masks(3 * (size_t)S, 0), act_full(S, 0);
  std::vector<double> sc((size_t)S * NSC);
  int it = 0, n64 = 0, ngraph = 0, nlo_it = 0;
  b->stat_launches = 0;
  b->stat_seconds = b->stat_bytes = 0;
  // by the precisions in use up to step M_lo/2 (bits 0-1: FP32, FP64) and past it (bits 2-3)
  GraphExecs<16> graphs;
  hipGraphExec_t* exec = graphs.e;
  bool warmed[16] = {};
  int result = LTB_OK;
  const int nsteps = (M + 1) / 2, nlo = two ? (prm->M_lo + 1) / 2 : nsteps;
  for (it = 0; it < prm->max_iter; it++) {
    int precs = 0, precs_full = 0;
    for (int s = 0; s < S; s++) {
      precs |= active[s] == 1 ? 1 : (active[s] == 2 ? 2 : 0);
      act_full[s] = (two && active[s] == 1 && low[s]) ? 0 : active[s];
      precs_full |= act_full[s] == 1 ? 1 : (act_full[s] == 2 ? 2 : 0);
    }
    if (!precs) break;
    if (!two) precs_full = 0;
    const int key = precs | (precs_full << 2);
    LTB_HIP(c, hipMemcpy(b->d_active, active.data(), sizeof(int) * S, hipMemcpyHostToDevice));
    if (two) {
      for (int s = 0; s < S; s++) {
        const bool two_s = active[s] == 1;  // FP32 structures run the two-level protocol
        masks[s] = act_full[s];
        masks[S + s] = two_s && low[s] ? 1 : 0;
        masks[2 * S + s] = two_s ? (low[s] ? 1 : 2) : 0;
      }
      LTB_HIP(c, hipMemcpy(b->d_masks, masks.data(), sizeof(int) * masks.size(), hipMemcpyHostToDevice));
      if (!precs_full) nlo_it++;
    }
    if (prm->use_graph && warmed[key]) {
      if (!exec[key]) {
        hipGraph_t g = nullptr;
        LTB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int s2 = batch_iteration(b, prm, precs, precs_full, true);
        const hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (s2) return s2;
        if (e != hipSuccess) return set_err(c, LTB_ERR_DEVICE, std::string("batch graph capture: ") + hipGetErrorString(e));
        const hipError_t e2 = hipGraphInstantiate(&exec[key], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) return set_err(c, LTB_ERR_DEVICE, "batch graph instantiate failed");
      }
      LTB_HIP(c, hipGraphLaunch(exec[key], c->stream));
      ngraph++;
    } else {
      const int s2 = batch_iteration(b, prm, precs, precs_full, false);
      if (s2) return s2;
      warmed[key] = true;
      if (c->timing) {
        LTB_HIP(c, hipStreamSynchronize(c->stream));
        const int nlast = (!two || precs_full) ? nsteps : nlo;
        for (int n = 0; n < nlast; n++) {
          float ms = 0;
          LTB_HIP(c, hipEventElapsedTime(&ms, b->events[2 * n], b->events[2 * n + 1]));
          b->stat_seconds += ms * 1e-3;
        }
        b->stat_launches += nlast;
        b->stat_bytes += iteration_bytes(b, active, 0, nlo) + iteration_bytes(b, act_full, nlo, nsteps);
      }
    }
    if (precs & 2) n64++;
    LTB_HIP(c, hipMemcpyAsync(sc.data(), b->d_sc, sizeof(double) * sc.size(), hipMemcpyDeviceToHost, c->stream));
    LTB_HIP(c, hipStreamSynchronize(c->stream));
    int n32 = 0, n64s = 0;
    for (int s = 0; s < S; s++) {
      if (!active[s]) continue;
      const double* p = sc.data() + (size_t)s * NSC;
      it_s[s]++;
      if (p[S_BAD] != 0.0) {
        st_s[s] = LTB_ERR_NONFINITE;
        active[s] = 0;
        result = LTB_ERR_NONFINITE;
        continue;
      }
      if (two && active[s] == 1) {
        if (low[s]) {  // the end of a low-order phase, as in ltb_scc_run
          plen[s]++;
          if (plen[s] == 1 || p[S_DQ] < 0.9 * best[s]) {
            best[s] = p[S_DQ];
            stall[s] = 0;
          } else {
            stall[s]++;
          }
          if (p[S_DQ] < lo_tol || p[S_DQ] < kTwoLevelPhaseRel * dq_full[s] || stall[s] >= kTwoLevelStall ||
              plen[s] >= kTwoLevelMaxPhase)
            low[s] = 0;
          n32++;
          continue;
        }
        dq_full[s] = p[S_DQ];
        if (p[S_DQ] < prm->tol && prm->switch_tol <= 0) {
          st_s[s] = LTB_OK;
          active[s] = 0;
          continue;
        }
        if (prm->switch_tol > 0 && p[S_DQ] < prm->switch_tol) {
          active[s] = 2;  // the plain FP64 finish
          n64s++;
        } else {
          low[s] = 1;
          plen[s] = 0;
          n32++;
        }
        continue;
      }
      // the switch rule of ltb_scc_run, per structure
      const bool final_prec = active[s] == 2 || prm->switch_tol <= 0;
      if (p[S_DQ] < prm->tol && final_prec) {
        st_s[s] = LTB_OK;
        active[s] = 0;
        continue;
      }
      if (active[s] == 1 && prm->switch_tol > 0 && p[S_DQ] < prm->switch_tol) active[s] = 2;
      (active[s] == 1 ? n32 : n64s)++;
    }
    if (prm->verbose) {
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      fprintf(stderr, "batch scc iter %d: %d fp32 + %d fp64 of %d structures still iterating, t=%.2fs\n", it,
              n32, n64s, S, el);
    }
  }
  if (p_out) LTB_HIP(c, hipMemcpy(p_out, b->d_pout, sizeof(double) * b->nat, hipMemcpyDeviceToHost));
  int nconv = 0;
  for (int s = 0; s < S; s++) {
    const double* p = sc.data() + (size_t)s * NSC;
    if (iters) iters[s] = it_s[s];
    if (eband_fp) eband_fp[s] = p[S_EBFP];
    if (mu_ha) mu_ha[s] = p[S_A] * p[S_MU] + p[S_B];
    if (status) status[s] = st_s[s];
    nconv += st_s[s] == LTB_OK;
  }
  if (total) {
    total->iterations = it;
    total->fp64_iterations = n64;
    total->dq = 0;
    total->mu_ha = 0;
    total->eband = 0;
    total->eband_fp = 0;
    total->graph_launches = ngraph;
    total->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    total->lo_iterations = nlo_it;
  }
  if (result) return set_err(c, result, "non-finite value in a batched structure");
  if (nconv < S) return set_err(c, LTB_NOT_CONVERGED, std::to_string(S - nconv) + " structures not converged");
  return LTB_OK;
}

int ltb_batch_set_tiling(ltb_batch* b, const ltb_kpe_tiling* t) {
  if (!b || !t) return LTB_ERR_ARG;
  if (t->atoms_per_wave < 0 || t->atoms_per_wave > 64 || t->blocks_per_cu < 0 || t->blocks_per_cu > 16)
    return set_err(b->ctx, LTB_ERR_ARG, "batch tiling: atoms_per_wave in 1..64, blocks_per_cu in 1..16");
  ltb_kpe_tiling n = b->tiling;
  if (t->atoms_per_wave) n.atoms_per_wave = t->atoms_per_wave;
  if (t->blocks_per_cu) n.blocks_per_cu = t->blocks_per_cu;
  if (t->nontemporal >= 0) n.nontemporal = t->nontemporal ? 1 : 0;
  b->tiling = n;
  return build_tasks(b);
}

int ltb_batch_get_tiling(const ltb_batch* b, ltb_kpe_tiling* out) {
  if (!b || !out) return LTB_ERR_ARG;
  *out = b->tiling;
  return LTB_OK;
}

int ltb_batch_autotune(ltb_batch* b, int precision, int trial_steps, ltb_kpe_tiling* best,
                       double* best_gbs) {
  if (!b || trial_steps < 1 || (precision != LTB_FP64 && precision != LTB_FP32_ACC64)) return LTB_ERR_ARG;
  ltb_ctx* c = b->ctx;
  int st = ensure_vectors(b);
  if (st) return st;
  // H~ of the neutral start for every structure
  std::vector<int> active(b->S, precision == LTB_FP64 ? 2 : 1);
  LTB_HIP(c, hipMemcpy(b->d_active, active.data(), sizeof(int) * b->S, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(b->d_pin, b->d_p0, sizeof(double) * b->nat, hipMemcpyDeviceToDevice));
  hipLaunchKernelGGL(k_batch_shift, dim3((b->nat + 255) / 256), dim3(256), 0, c->stream, b->d_grow,
                     b->d_gcol, b->d_gval, b->d_pin, b->d_p0, b->d_atomst, b->d_active, b->d_v, b->nat);
  hipLaunchKernelGGL(k_batch_gersh, dim3(b->S), dim3(256), 0, c->stream, b->d_diag0, b->d_offabs, b->d_v,
                     b->d_aoff, b->d_active, b->d_sc);
  hipLaunchKernelGGL(k_batch_rescale, dim3((b->nnzb * 16 + 255) / 256), dim3(256), 0, c->stream, b->d_h0,
                     b->d_blkrow, b->d_col, b->d_atomst, b->d_v, b->d_sc, b->d_active, b->nnzb, b->d_ht64,
                     b->d_ht32);
  LTB_HIP(c, hipGetLastError());
  // bytes of one step n >= 2 over all structures
  double step_bytes = 0;
  const double el = precision == LTB_FP64 ? 8.0 : 4.0;
  for (int s = 0; s < b->S; s++) {
    const double norb = 4.0 * (b->aoff[s + 1] - b->aoff[s]);
    step_bytes += 3 * el * norb * norb + b->nnzb_s[s] * (16.0 * el + 4.0) + 4.0 * (norb / 4 + 1);
  }
  hipEvent_t e0, e1;
  LTB_HIP(c, hipEventCreate(&e0));
  LTB_HIP(c, hipEventCreate(&e1));
  const ltb_kpe_tiling keep = b->tiling;
  ltb_kpe_tiling top = keep;
  double top_gbs = -1;
  // one flat list of candidates, so the first failure leaves the search
  struct Cand {
    int apw, bpc, nt;
  };
  std::vector<Cand> cands;
  for (int apw : {4, 8, 16})
    for (int bpc : {2, 4, 8})
      for (int nt : {0, 1}) cands.push_back({apw, bpc, nt});
  for (const Cand& cd : cands) {
    b->tiling.atoms_per_wave = cd.apw;
    b->tiling.blocks_per_cu = cd.bpc;
    b->tiling.nontemporal = cd.nt;
    if ((st = build_tasks(b))) break;
    const int grid = step_grid(b);
    auto step = [&](int n) {
      const int mode = n == 0 ? 0 : (n == 1 ? 1 : 2);
      void *cur, *nxt;
      if (precision == LTB_FP64) {
        step_buffers<double>(b, n, &cur, &nxt);
        launch_step<double>(b, mode, grid, cur, nxt);
      } else {
        step_buffers<float>(b, n, &cur, &nxt);
        launch_step<float>(b, mode, grid, cur, nxt);
      }
    };
    step(0);
    step(1);
    (void)hipEventRecord(e0, c->stream);
    for (int n = 2; n < 2 + trial_steps; n++) step(n);
    (void)hipEventRecord(e1, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess || hipGetLastError() != hipSuccess) {
      st = set_err(c, LTB_ERR_DEVICE, "batch autotune trial failed");
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double gbs = step_bytes * trial_steps / (ms * 1e-3) / 1e9;
    if (gbs > top_gbs) {
      top_gbs = gbs;
      top = b->tiling;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st) {
    b->tiling = keep;
    (void)build_tasks(b);
    return st;
  }
  b->tiling = top;
  if ((st = build_tasks(b))) return st;
  if (best) *best = top;
  if (best_gbs) *best_gbs = top_gbs;
  return LTB_OK;
}

}  // extern "C"
