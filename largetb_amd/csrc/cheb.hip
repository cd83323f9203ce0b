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
// and W = 64*V columns (a "column group"); every lane holds V consecutive
// columns, so each row access of a wave is one contiguous W*sizeof(T)
// segment.  Block-row pointers, column indices and the 4x4 blocks of H~ are
// wave-uniform (scalar loads).  The rows of up to 4 neighbour blocks and the
// t_{n-1} rows are all requested before the first FMA.  t_{n-1} and t_{n+1}
// are touched once per step and may bypass the caches (non-temporal); the
// t_n rows are re-read by every neighbouring atom and should hit L2.
//
// Two schedules:
//  * blocked: workgroup (column group, 4*apw consecutive atoms), ids remapped
//    so each XCD gets a contiguous range;
//  * sweep: a resident grid of 8*slots workgroups; workgroup b runs on XCD
//    b%8 and that XCD owns column groups xcd, xcd+8, ...; its waves sweep the
//    atoms of one column group cyclically (wave w takes atoms w, w+S, ...), so
//    the rows in flight on an XCD form one compact window and the t_n rows a
//    neighbour re-reads (a few hundred to ~1300 atoms away) are still in that
//    XCD's L2.
// Per-column dot partials are summed in LDS per workgroup, written once, and
// folded into per-atom moments (fold_columns): under the sweep schedule by the
// first workgroups of the next step's launch, otherwise (and after a chunk's
// last step) by k_cheb_reduce.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "cheb_device.hpp"
#include "ctx.hpp"
#include "system.hpp"

namespace {

using ltbdev::atom_step;
using ltbdev::flush_partials;

// Fold the workgroup partials of step n into per-atom moments:
//   n == 0: mu_1 = <t_1, t_0>,           mu_2 = 2 <t_1, t_1> - 1
//   n >= 1: mu_{2n+1} = 2 <t_{n+1}, t_n> - mu_1,  mu_{2n+2} = 2 |t_{n+1}|^2 - 1
// and sum the 4 columns (orbitals) of each atom.  mom is [M+1][nat].  The
// partial rows holding column c are rows [base(c), base(c) + count): blocked
// schedule base 0, count n_sg; sweep schedule base (c/W % 8)*slots, count slots.
struct FoldArgs {
  const double* part;  // partial rows of the step being folded (nullptr: none)
  int count, sweep_W, slots, ld, nvalid, j0, n, M;
  double* mu1col;
  double* mom;
  int nat;
  double* momorb;  // per-orbital moments [M+1][norb], or nullptr
  int norb;
};

// the 256 threads of a workgroup fold columns 64 cb .. 64 cb + 63
__device__ __forceinline__ void fold_columns(const FoldArgs& f, int cb) {
  const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const int c = cb * 64 + lane;
  const int base = f.sweep_W > 0 ? ((c / f.sweep_W) & 7) * f.slots : 0;
  double sxy = 0, syy = 0;
  for (int s = slice; s < f.count; s += 4) {
    sxy += f.part[(size_t)(2 * (base + s)) * f.ld + c];
    syy += f.part[(size_t)(2 * (base + s) + 1) * f.ld + c];
  }
  __shared__ double sh[2][4][64];
  sh[0][slice][lane] = sxy;
  sh[1][slice][lane] = syy;
  __syncthreads();
  if (slice == 0) {
    sxy = (sh[0][0][lane] + sh[0][1][lane]) + (sh[0][2][lane] + sh[0][3][lane]);
    syy = (sh[1][0][lane] + sh[1][1][lane]) + (sh[1][2][lane] + sh[1][3][lane]);
    const bool valid = c < f.nvalid;
    const int n = f.n, M = f.M;
    double modd;
    if (n == 0) {
      modd = sxy;
      f.mu1col[c] = sxy;
    } else {
      modd = 2.0 * sxy - f.mu1col[c];
    }
    double meven = 2.0 * syy - (valid ? 1.0 : 0.0);
    if (f.momorb && valid) {  // per-orbital moments [M+1][norb] (diag(P) per orbital)
      const int j = f.j0 + c;
      if (n == 0) f.momorb[j] = 1.0;
      if (2 * n + 1 <= M) f.momorb[(size_t)(2 * n + 1) * f.norb + j] = modd;
      if (2 * n + 2 <= M) f.momorb[(size_t)(2 * n + 2) * f.norb + j] = meven;
    }
    modd += __shfl_xor(modd, 1);
    modd += __shfl_xor(modd, 2);
    meven += __shfl_xor(meven, 1);
    meven += __shfl_xor(meven, 2);
    if ((lane & 3) == 0 && valid) {
      const int atom = (f.j0 + c) >> 2;
      if (n == 0) f.mom[atom] = 4.0;
      const int mo = 2 * n + 1, me = 2 * n + 2;
      if (mo <= M) f.mom[(size_t)mo * f.nat + atom] = modd;
      if (me <= M) f.mom[(size_t)me * f.nat + atom] = meven;
    }
  }
  __syncthreads();  // sh is free again
}

template <typename T, int V, int MODE, bool NT>
__global__ __launch_bounds__(256) void k_cheb_blocked(const int* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const T* __restrict__ hv,
                                                      const T* __restrict__ xcur,
                                                      T* __restrict__ xnext, int nat, int ld,
                                                      int j0, int nvalid, int apw, int n_sg,
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
  double sxy[V], syy[V];
#pragma unroll
  for (int v = 0; v < V; v++) sxy[v] = syy[v] = 0.0;
  for (int A = a_begin; A < a_end; ++A)
    atom_step<T, V, MODE, NT>(rowptr, col, hv, xcur, xnext, A, ld, j0, c0, nvalid, sxy, syy);
  flush_partials<V>(sxy, syy, wave, lane, part, (size_t)sg * 2, ld, cg * W);
}

template <typename T, int V, int MODE, bool NT>
__global__ __launch_bounds__(256) void k_cheb_sweep(const int* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const T* __restrict__ hv,
                                                    const T* __restrict__ xcur,
                                                    T* __restrict__ xnext, int nat, int ld, int j0,
                                                    int nvalid, int n_cg, int slots,
                                                    const int* __restrict__ visit,
                                                    double* __restrict__ part, FoldArgs prev,
                                                    const int* __restrict__ ranges, int nranges) {
  constexpr int W = 64 * V;
  // the previous step's partials (complete: it was the previous launch) are
  // folded here instead of in a launch of their own
  if (prev.part)
    for (int cb = blockIdx.x; cb < prev.ld / 64; cb += gridDim.x) fold_columns(prev, cb);
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int S = 4 * slots;
  const int wid = 4 * slot + wave;
  for (int cg = xcd; cg < n_cg; cg += 8) {
    const int c0 = cg * W + lane * V;
    double sxy[V], syy[V];
#pragma unroll
    for (int v = 0; v < V; v++) sxy[v] = syy[v] = 0.0;
    if (!ranges) {
      for (int i = wid; i < nat; i += S)
        atom_step<T, V, MODE, NT>(rowptr, col, hv, xcur, xnext, visit ? visit[i] : i, ld, j0, c0, nvalid, sxy,
                                  syy);
    } else {
      // the chunk's light cone: cyclic atom ranges {lo, len}, swept as one
      // sequence (wave w takes positions w, w + S, ... of their concatenation)
      int base = 0;
      for (int r = 0; r < nranges; r++) {
        const int lo = ranges[2 * r], len = ranges[2 * r + 1];
        int i = (wid - base) % S;
        i += i < 0 ? S : 0;
        for (; i < len; i += S) {
          int A = lo + i;
          A -= A >= nat ? nat : 0;
          atom_step<T, V, MODE, NT>(rowptr, col, hv, xcur, xnext, A, ld, j0, c0, nvalid, sxy, syy);
        }
        base += len;
      }
    }
    flush_partials<V>(sxy, syy, wave, lane, part, (size_t)(xcd * slots + slot) * 2, ld, cg * W);
  }
}

// the partials of the last step of a chunk (and of every step under the
// blocked schedule)
__global__ __launch_bounds__(256) void k_cheb_reduce(FoldArgs f) { fold_columns(f, blockIdx.x); }

// Zero the 4 rows (ld elements each, contiguous) of every atom of the cyclic
// ranges {lo, len} (count = the ranges' total length), or of atoms 0 .. count-1
// without ranges.  Workgroup i takes the i-th atom of the concatenated ranges.
template <typename T>
__global__ __launch_bounds__(256) void k_zero_atoms(T* __restrict__ buf, const int* __restrict__ ranges, int nranges,
                                                    int nat, int ld, int count) {
  const int n16 = (int)(4 * (size_t)ld * sizeof(T) / 16);
  for (int i = blockIdx.x; i < count; i += gridDim.x) {
    int A = i;
    if (ranges) {
      int base = 0, r = 0;
      while (r < nranges - 1 && i >= base + ranges[2 * r + 1]) base += ranges[2 * r + 1], r++;
      A = ranges[2 * r] + (i - base);
      A -= A >= nat ? nat : 0;
    }
    float4* row = reinterpret_cast<float4*>(buf + (size_t)4 * A * ld);
    for (int k = threadIdx.x; k < n16; k += 256) row[k] = make_float4(0.f, 0.f, 0.f, 0.f);
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

// Per-column moments of explicit start vectors t_0 = r_c (stochastic trace,
// getKPE on given vectors):  mu_0 = |r_c|^2,  mu_1 = <t_1, t_0>,
//   mu_{2n+1} = 2 <t_{n+1}, t_n> - mu_1,  mu_{2n+2} = 2 |t_{n+1}|^2 - |r_c|^2
// i.e. r_c^T T_m r_c; momcol is [M+1][ld].  Partial rows as k_cheb_reduce.
__global__ __launch_bounds__(256) void k_block_reduce(const double* __restrict__ part, int count,
                                                      int sweep_W, int slots, int ld, int nvalid, int n,
                                                      int M, double* __restrict__ mu1col,
                                                      const double* __restrict__ r2col,
                                                      double* __restrict__ momcol) {
  const int lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + lane;
  const int base = sweep_W > 0 ? ((c / sweep_W) & 7) * slots : 0;
  double sxy = 0, syy = 0;
  for (int s = slice; s < count; s += 4) {
    sxy += part[(size_t)(2 * (base + s)) * ld + c];
    syy += part[(size_t)(2 * (base + s) + 1) * ld + c];
  }
  __shared__ double sh[2][4][64];
  sh[0][slice][lane] = sxy;
  sh[1][slice][lane] = syy;
  __syncthreads();
  if (slice != 0 || c >= nvalid) return;
  sxy = (sh[0][0][lane] + sh[0][1][lane]) + (sh[0][2][lane] + sh[0][3][lane]);
  syy = (sh[1][0][lane] + sh[1][1][lane]) + (sh[1][2][lane] + sh[1][3][lane]);
  const double r2 = r2col[c];
  double modd;
  if (n == 0) {
    modd = sxy;
    mu1col[c] = sxy;
    momcol[c] = r2;
  } else {
    modd = 2.0 * sxy - mu1col[c];
  }
  if (2 * n + 1 <= M) momcol[(size_t)(2 * n + 1) * ld + c] = modd;
  if (2 * n + 2 <= M) momcol[(size_t)(2 * n + 2) * ld + c] = 2.0 * syy - r2;
}

// atoms a launch of the sweep schedule visits: n cyclic ranges {lo, len} at
// d (device, 2 ints each); d == nullptr: all of them
struct AtomRange {
  const int* d = nullptr;
  int n = 0;
};

struct Plan {
  int V, W, ld, schedule, nt, order;
  int apw, n_sg;   // blocked
  int slots;       // sweep
  size_t part_rows;
};

template <typename T>
Plan make_plan(const ltb_ctx* c) {
  const KpeTiling& t = c->tiling[sizeof(T) == 8 ? 1 : 0];
  Plan p{};
  const bool big = c->norb >= 4096;
  p.V = t.cols_per_lane > 0 ? t.cols_per_lane : (big ? (sizeof(T) == 4 ? 2 : 1) : 1);
  if (p.V != 1 && p.V != 2 && p.V != 4) p.V = 1;
  p.W = 64 * p.V;
  int want = t.chunk_cols > 0 ? t.chunk_cols : (sizeof(T) == 4 ? 4096 : 2048);
  int ld = std::min(want, c->norb);
  p.ld = (ld + p.W - 1) / p.W * p.W;
  const int n_cg = p.ld / p.W;
  p.schedule = t.schedule;
  if (p.schedule == LTB_SCHED_AUTO) p.schedule = (n_cg >= 16 && big) ? LTB_SCHED_SWEEP : LTB_SCHED_BLOCKED;
  if (p.schedule == LTB_SCHED_SWEEP && n_cg < 8) p.schedule = LTB_SCHED_BLOCKED;
  p.nt = t.nontemporal < 0 ? 1 : t.nontemporal;
  p.order = t.order == 1 ? 1 : 2;
  const int cus = std::max(c->cu_count, 8);
  if (p.schedule == LTB_SCHED_SWEEP) {
    int bpc = t.blocks_per_cu > 0 ? t.blocks_per_cu : 4;
    p.slots = std::max(1, bpc * cus / 8);
    p.part_rows = (size_t)8 * p.slots;
  } else {
    if (t.atoms_per_wave > 0) {
      p.apw = t.atoms_per_wave;
    } else {
      int target_blocks = 8 * cus;
      int sg_want = std::max(1, (target_blocks + n_cg - 1) / n_cg);
      p.apw = std::max(1, (c->nat + 4 * sg_want - 1) / (4 * sg_want));
    }
    p.n_sg = (c->nat + 4 * p.apw - 1) / (4 * p.apw);
    p.part_rows = p.n_sg;
  }
  return p;
}

// part: where this step's partials go; prev: the previous step's partials to
// fold first (sweep schedule only; prev.part == nullptr for none)
template <typename T, int V, int MODE, bool NT>
void launch_one(const ltb_ctx* c, const T* hv, const T* xcur, T* xnext, int ldc, int j0, int nvalid,
                const Plan& p, double* part, const FoldArgs& prev, AtomRange r) {
  const int n_cg = ldc / (64 * V);
  if (p.schedule == LTB_SCHED_SWEEP) {
    hipLaunchKernelGGL((k_cheb_sweep<T, V, MODE, NT>), dim3(8 * p.slots), dim3(256), 0, c->stream,
                       c->d_rowptr, c->d_col, hv, xcur, xnext, c->nat, ldc, j0, nvalid, n_cg,
                       p.slots, p.order == 1 ? c->d_visit : nullptr, part, prev, r.d, r.n);
  } else {
    hipLaunchKernelGGL((k_cheb_blocked<T, V, MODE, NT>), dim3(n_cg * p.n_sg), dim3(256), 0,
                       c->stream, c->d_rowptr, c->d_col, hv, xcur, xnext, c->nat, ldc, j0, nvalid,
                       p.apw, p.n_sg, part);
  }
}

template <typename T, int V, int MODE>
void launch_nt(const ltb_ctx* c, const T* hv, const T* xcur, T* xnext, int ldc, int j0, int nvalid,
               const Plan& p, double* part, const FoldArgs& prev, AtomRange r) {
  if (p.nt)
    launch_one<T, V, MODE, true>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else
    launch_one<T, V, MODE, false>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
}

template <typename T, int V>
void launch_mode(const ltb_ctx* c, int mode, const T* hv, const T* xcur, T* xnext, int ldc, int j0,
                 int nvalid, const Plan& p, double* part, const FoldArgs& prev, AtomRange r) {
  if (mode == 0)
    launch_nt<T, V, 0>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else if (mode == 1)
    launch_nt<T, V, 1>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else if (mode == 2)
    launch_nt<T, V, 2>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else
    launch_nt<T, V, 3>(c, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
}

template <typename T>
void launch_step(const ltb_ctx* c, int mode, const T* hv, const T* xcur, T* xnext, int ldc, int j0,
                 int nvalid, const Plan& p, double* part, const FoldArgs& prev = FoldArgs{},
                 AtomRange r = AtomRange{}) {
  if (p.V == 4)
    launch_mode<T, 4>(c, mode, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else if (p.V == 2)
    launch_mode<T, 2>(c, mode, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
  else
    launch_mode<T, 1>(c, mode, hv, xcur, xnext, ldc, j0, nvalid, p, part, prev, r);
}

// Light cones of the unit-vector chunks.  t_{n+1} = T_{n+1}(H~) e_j is zero
// on every atom more than n + 1 hops (in H~'s block graph) from atom j/4, so
// step n of the chunk j0 .. j0 + ld - 1 only has to visit the atoms within
// n + 1 hops of its own.  Both buffers start the chunk all zero (see
// run_moments).  A row a step leaves alone is then outside that step's cone
// and so outside every earlier one: each value the buffers ever held there is
// an exact zero (a product or sum of zeros), which is what its readers need,
// and a chunk leaves non-zeros only inside its last cone.  Visiting a
// superset of the cone is therefore always safe; per chunk and hop count h
// the table holds the cone as at most kMaxRanges cyclic atom ranges {lo, len}
// (runs of consecutive atoms, the closest runs merged), or {0, nat} once it
// holds every atom.
constexpr int kMaxRanges = 32;

void free_cones_impl(ltb_ctx* c) {
  for (auto& kv : c->cones)
    if (kv.second.d_ranges) (void)hipFree(kv.second.d_ranges);
  c->cones.clear();
}

int build_cones(ltb_ctx* c, int ld) {
  if (c->cones.count(ld)) return LTB_OK;
  const int nat = c->nat;
  if ((int)c->h_rowptr.size() != nat + 1) {  // kept only once both copies succeeded
    std::vector<int> rp(nat + 1), cl(c->nnzb);
    LTB_HIP(c, hipMemcpy(rp.data(), c->d_rowptr, sizeof(int) * (nat + 1), hipMemcpyDeviceToHost));
    LTB_HIP(c, hipMemcpy(cl.data(), c->d_col, sizeof(int) * c->nnzb, hipMemcpyDeviceToHost));
    c->h_rowptr.swap(rp);
    c->h_col.swap(cl);
  }
  ltb_ctx::ConeTable t;
  ltb::light_cones(nat, c->h_rowptr.data(), c->h_col.data(), ld / 4, kMaxRanges, t.idx, t.ranges);
  std::vector<int> flat;
  for (const auto& r : t.ranges) {
    flat.push_back(r.first);
    flat.push_back(r.second);
  }
  LTB_HIP(c, hipMalloc(&t.d_ranges, sizeof(int) * std::max<size_t>(flat.size(), 2)));
  hipError_t e = hipMemcpy(t.d_ranges, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(t.d_ranges);
    return set_err(c, LTB_ERR_DEVICE, std::string("light-cone upload: ") + hipGetErrorString(e));
  }
  c->cones[ld] = std::move(t);
  return LTB_OK;
}

template <typename T>
int ensure_buffers(ltb_ctx* c, const Plan& p, int M) {
  size_t xb = (size_t)c->norb * p.ld * sizeof(T);
  int st;
  if ((st = ensure_buffer(c, &c->d_xa, &c->xabytes, xb))) return st;
  if ((st = ensure_buffer(c, &c->d_xb, &c->xbbytes, xb))) return st;
  // two sets of partial rows: step n writes set n % 2 while it folds set (n - 1) % 2
  if ((st = ensure_buffer(c, (void**)&c->d_part, &c->partbytes, 2 * p.part_rows * 2 * p.ld * sizeof(double))))
    return st;
  if ((st = ensure_buffer(c, (void**)&c->d_mu1, &c->mu1bytes, (size_t)p.ld * sizeof(double)))) return st;
  if (M >= 0) {
    if ((st = ensure_buffer(c, (void**)&c->d_mom, &c->mombytes,
                            (size_t)(M + 1) * c->nat * sizeof(double) + (size_t)(M + 1) * sizeof(double))))
      return st;
    c->d_momtot = c->d_mom + (size_t)(M + 1) * c->nat;
    if (c->keep_orbital &&
        (st = ensure_buffer(c, (void**)&c->d_momorb, &c->momorbbytes, (size_t)(M + 1) * c->norb * sizeof(double))))
      return st;
  }
  return LTB_OK;
}

template <typename T>
int run_moments(ltb_ctx* c, int M) {
  const Plan p = make_plan<T>(c);
  const int nsteps = (M + 1) / 2;  // step n yields moments up to 2n+2
  int st = ensure_buffers<T>(c, p, M);
  if (st) return st;
  const T* hv = sizeof(T) == 4 ? (const T*)c->d_ht32 : (const T*)c->d_ht64;
  T* A = (T*)c->d_xa;
  T* B = (T*)c->d_xb;
  const int nlaunch_est = ((c->norb + p.ld - 1) / p.ld) * nsteps;
  if (c->timing) {
    while ((int)c->events.size() < 2 * nlaunch_est) {
      hipEvent_t e;
      LTB_HIP(c, hipEventCreate(&e));
      c->events.push_back(e);
    }
  }
  const double hbytes = (double)c->nnzb * (16.0 * sizeof(T) + 4.0) + 4.0 * (c->nat + 1);
  // light cones only matter with more than one chunk, on the natural order
  const bool cones = c->light_cone && p.schedule == LTB_SCHED_SWEEP && p.order == 2 && c->norb > p.ld;
  if (cones && (st = build_cones(c, p.ld))) return st;
  int launches = 0;
  double bytes = 0;
  // With light cones every step of a chunk visits only its cone, so the
  // buffers must hold zeros wherever a step reads rows it has not written:
  // both are cleared before the first chunk, and after a chunk only its last
  // cone (the one place it left anything but zeros) is cleared again.
  bool clear = true;
  const int* dirty = nullptr;  // device ranges of the previous chunk's last cone
  int ndirty = 0, dirty_atoms = 0, dirty_ldc = 0;
  auto zero_atoms = [&](T* buf, const int* d_ranges, int nranges, int count, int ldc) {
    if (count > 0)
      hipLaunchKernelGGL(k_zero_atoms<T>, dim3(std::min(count, 65535)), dim3(256), 0, c->stream, buf, d_ranges,
                         nranges, c->nat, ldc, count);
  };
  for (int j0 = 0; j0 < c->norb; j0 += p.ld) {
    const int nvalid = std::min(p.ld, c->norb - j0);
    const int ldc = (nvalid + p.W - 1) / p.W * p.W;
    const int n_cg = ldc / p.W;
    // a narrow last chunk may leave XCDs idle under the sweep schedule
    Plan pc = p;
    if (pc.schedule == LTB_SCHED_SWEEP && n_cg < 8) {
      pc.schedule = LTB_SCHED_BLOCKED;
      int target_blocks = 8 * std::max(c->cu_count, 8);
      int sg_want = std::max(1, (target_blocks + n_cg - 1) / n_cg);
      pc.apw = std::max(1, (c->nat + 4 * sg_want - 1) / (4 * sg_want));
      pc.n_sg = (c->nat + 4 * pc.apw - 1) / (4 * pc.apw);
      if ((size_t)pc.n_sg > p.part_rows) {
        pc.n_sg = (int)p.part_rows;
        pc.apw = (c->nat + 4 * pc.n_sg - 1) / (4 * pc.n_sg);
        pc.n_sg = (c->nat + 4 * pc.apw - 1) / (4 * pc.apw);
      }
    }
    const bool sweep = pc.schedule == LTB_SCHED_SWEEP;
    const size_t half = p.part_rows * 2 * (size_t)p.ld;
    auto fold_args = [&](int n) {
      return FoldArgs{c->d_part + (n & 1) * half, sweep ? pc.slots : pc.n_sg, sweep ? pc.W : 0, pc.slots,
                      ldc, nvalid, j0, n, M, c->d_mu1, c->d_mom, c->nat,
                      c->keep_orbital ? c->d_momorb : nullptr, c->norb};
    };
    const bool chunk_cones = cones && sweep;
    if (chunk_cones) {
      if (clear || ldc != dirty_ldc) {  // a narrower last chunk lays its rows out at its own width
        zero_atoms(A, nullptr, 0, c->nat, ldc);
        zero_atoms(B, nullptr, 0, c->nat, ldc);
      } else if (ndirty > 0) {
        zero_atoms(A, dirty, ndirty, dirty_atoms, ldc);
        zero_atoms(B, dirty, ndirty, dirty_atoms, ldc);
      }
      LTB_HIP(c, hipGetLastError());
      const auto& t = c->cones[p.ld];
      const auto& ci = t.idx[j0 / p.ld];
      const auto& e = ci[std::min<size_t>(nsteps, ci.size() - 1)];  // holds t_{nsteps-1} and t_{nsteps}
      dirty = t.d_ranges + 2 * (size_t)e.first;
      ndirty = e.second;
      dirty_atoms = 0;
      for (int q = 0; q < e.second; q++) dirty_atoms += t.ranges[e.first + q].second;
      dirty_ldc = ldc;
      clear = false;
    } else {
      clear = true;  // a chunk without cones writes every row at its own width
    }
    for (int n = 0; n < nsteps; n++) {
      const int mode = n == 0 ? 0 : (n == 1 ? 1 : 2);
      T* cur = (n & 1) ? B : A;
      T* nxt = (n == 0) ? B : ((n & 1) ? A : B);
      // under the sweep schedule step n folds step n - 1's partials itself
      const FoldArgs prev = sweep && n > 0 ? fold_args(n - 1) : FoldArgs{};
      // step n visits the cone of t_{n+1} (every row without light cones)
      AtomRange r;
      const std::pair<int, int>* rr = nullptr;
      if (chunk_cones) {
        const auto& t = c->cones[p.ld];
        const auto& ci = t.idx[j0 / p.ld];
        const auto& e = ci[std::min<size_t>(n + 1, ci.size() - 1)];
        rr = t.ranges.data() + e.first;
        if (!(e.second == 1 && rr->second >= c->nat)) {
          r.d = t.d_ranges + 2 * (size_t)e.first;
          r.n = e.second;
        }
      }
      if (c->timing) LTB_HIP(c, hipEventRecord(c->events[2 * launches], c->stream));
      launch_step<T>(c, mode, hv, cur, nxt, ldc, j0, nvalid, pc, c->d_part + (n & 1) * half, prev, r);
      if (c->timing) LTB_HIP(c, hipEventRecord(c->events[2 * launches + 1], c->stream));
      LTB_HIP(c, hipGetLastError());
      const int streams = mode == 0 ? 1 : (mode == 1 ? 2 : 3);
      if (!r.d) {
        bytes += (double)streams * sizeof(T) * c->norb * (double)nvalid + hbytes;
      } else {  // the rows and blocks of the visited atoms only
        const int* rp = c->h_rowptr.data();
        for (int q = 0; q < r.n; q++) {
          const int lo = rr[q].first, len = rr[q].second, e = lo + len;
          const double nb = e <= c->nat ? rp[e] - rp[lo] : (rp[c->nat] - rp[lo]) + rp[e - c->nat];
          bytes += (double)streams * sizeof(T) * 4.0 * len * (double)nvalid + nb * (16.0 * sizeof(T) + 4.0) +
                   4.0 * (len + 1);
        }
      }
      launches++;
      if (!sweep || n == nsteps - 1) {
        hipLaunchKernelGGL(k_cheb_reduce, dim3(ldc / 64), dim3(256), 0, c->stream, fold_args(n));
        LTB_HIP(c, hipGetLastError());
      }
    }
  }
  hipLaunchKernelGGL(k_mom_total, dim3(M + 1), dim3(256), 0, c->stream, c->d_mom, c->nat,
                     c->d_momtot);
  LTB_HIP(c, hipGetLastError());
  if (c->timing) {  // the statistics stay those of the last timed call
    LTB_HIP(c, hipStreamSynchronize(c->stream));
    double tot = 0;
    for (int i = 0; i < launches; i++) {
      float ms = 0;
      LTB_HIP(c, hipEventElapsedTime(&ms, c->events[2 * i], c->events[2 * i + 1]));
      tot += ms * 1e-3;
    }
    c->stat_launches = launches;
    c->stat_bytes = bytes;
    c->stat_seconds = tot;
  }
  c->mom_M = M;
  return LTB_OK;
}

// time `steps` general recursion steps (MODE 2) of one full-width chunk
template <typename T>
int time_steps(ltb_ctx* c, int steps, double* seconds, double* bytes) {
  const Plan p = make_plan<T>(c);
  int st = ensure_buffers<T>(c, p, -1);
  if (st) return st;
  const T* hv = sizeof(T) == 4 ? (const T*)c->d_ht32 : (const T*)c->d_ht64;
  const int nvalid = std::min(p.ld, c->norb);
  const int ldc = (nvalid + p.W - 1) / p.W * p.W;
  LTB_HIP(c, hipMemsetAsync(c->d_xa, 0, (size_t)c->norb * ldc * sizeof(T), c->stream));
  LTB_HIP(c, hipMemsetAsync(c->d_xb, 0, (size_t)c->norb * ldc * sizeof(T), c->stream));
  hipEvent_t e0, e1;
  LTB_HIP(c, hipEventCreate(&e0));
  LTB_HIP(c, hipEventCreate(&e1));
  launch_step<T>(c, 2, hv, (T*)c->d_xa, (T*)c->d_xb, ldc, 0, nvalid, p, c->d_part);  // warm
  LTB_HIP(c, hipEventRecord(e0, c->stream));
  for (int i = 0; i < steps; i++)
    launch_step<T>(c, 2, hv, (i & 1) ? (T*)c->d_xb : (T*)c->d_xa, (i & 1) ? (T*)c->d_xa : (T*)c->d_xb,
                   ldc, 0, nvalid, p, c->d_part);
  LTB_HIP(c, hipEventRecord(e1, c->stream));
  LTB_HIP(c, hipEventSynchronize(e1));
  LTB_HIP(c, hipGetLastError());
  float ms = 0;
  LTB_HIP(c, hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *seconds = ms * 1e-3;
  const double hbytes = (double)c->nnzb * (16.0 * sizeof(T) + 4.0) + 4.0 * (c->nat + 1);
  *bytes = steps * (3.0 * sizeof(T) * c->norb * (double)nvalid + hbytes);
  return LTB_OK;
}

// Moments r_c^T T_m(H~) r_c, m = 0..M, of nvec explicit vectors (columns of
// R, row-major [norb][ld] already in d_xa as T, zero beyond nvec) by the
// same recursion with moment doubling: step 0 is MODE 3 (t_1 = H~ t_0 with
// t_0 read), then MODE 2 over the ping-pong pair (t_0 stays in d_xa, so
// step 1 reads it as t_{n-1}).  One chunk of ld = nvec rounded up.
template <typename T>
int run_block_moments(ltb_ctx* c, int M, int nvec, const double* d_r2, double* d_momcol, int ld) {
  Plan p = make_plan<T>(c);
  p.ld = ld;
  const int n_cg = ld / p.W;
  if (p.schedule == LTB_SCHED_SWEEP && n_cg < 8) p.schedule = LTB_SCHED_BLOCKED;
  if (p.schedule == LTB_SCHED_BLOCKED) {
    const int target_blocks = 8 * std::max(c->cu_count, 8);
    const int sg_want = std::max(1, (target_blocks + n_cg - 1) / n_cg);
    p.apw = std::max(1, (c->nat + 4 * sg_want - 1) / (4 * sg_want));
    p.n_sg = (c->nat + 4 * p.apw - 1) / (4 * p.apw);
    p.part_rows = p.n_sg;
  }
  int st;
  if ((st = ensure_buffer(c, (void**)&c->d_part, &c->partbytes, p.part_rows * 2 * (size_t)ld * sizeof(double))))
    return st;
  if ((st = ensure_buffer(c, (void**)&c->d_mu1, &c->mu1bytes, (size_t)ld * sizeof(double)))) return st;
  const T* hv = sizeof(T) == 4 ? (const T*)c->d_ht32 : (const T*)c->d_ht64;
  T* A = (T*)c->d_xa;
  T* B = (T*)c->d_xb;
  const bool sweep = p.schedule == LTB_SCHED_SWEEP;
  const int nsteps = (M + 1) / 2;
  for (int n = 0; n < nsteps; n++) {
    const int mode = n == 0 ? 3 : 2;
    T* cur = (n & 1) ? B : A;
    T* nxt = (n & 1) ? A : B;
    launch_step<T>(c, mode, hv, cur, nxt, ld, 0, nvec, p, c->d_part);
    LTB_HIP(c, hipGetLastError());
    hipLaunchKernelGGL(k_block_reduce, dim3(ld / 64), dim3(256), 0, c->stream, c->d_part,
                       sweep ? p.slots : p.n_sg, sweep ? p.W : 0, p.slots, ld, nvec, n, M, c->d_mu1, d_r2,
                       d_momcol);
    LTB_HIP(c, hipGetLastError());
  }
  return LTB_OK;
}

}  // namespace

int kpe_block_moments_impl(ltb_ctx* c, int M, int nvec, int precision, const double* d_r2, double* d_momcol,
                           int ld) {
  if (precision == LTB_FP32_ACC64) return run_block_moments<float>(c, M, nvec, d_r2, d_momcol, ld);
  return run_block_moments<double>(c, M, nvec, d_r2, d_momcol, ld);
}

int kpe_block_width(ltb_ctx* c, int precision, int nvec) {
  const int W = precision == LTB_FP32_ACC64 ? make_plan<float>(c).W : make_plan<double>(c).W;
  return (nvec + W - 1) / W * W;
}

int kpe_moments_impl(ltb_ctx* c, int M, int precision) {
  if (precision == LTB_FP32_ACC64) return run_moments<float>(c, M);
  return run_moments<double>(c, M);
}

void free_cones(ltb_ctx* c) { free_cones_impl(c); }

int kpe_reserve(ltb_ctx* c, int M, int precision) {
  if (precision == LTB_FP32_ACC64) return ensure_buffers<float>(c, make_plan<float>(c), M);
  return ensure_buffers<double>(c, make_plan<double>(c), M);
}

int kpe_time_steps(ltb_ctx* c, int precision, int steps, double* seconds, double* bytes) {
  if (precision == LTB_FP32_ACC64) return time_steps<float>(c, steps, seconds, bytes);
  return time_steps<double>(c, steps, seconds, bytes);
}

int kpe_time_recursion(ltb_ctx* c, int precision, int M, int cols, int reps, double* seconds) {
  KpeTiling& t = c->tiling[precision == LTB_FP64 ? 1 : 0];
  const int saved_cols = t.chunk_cols;
  const bool saved_timing = c->timing;
  t.chunk_cols = cols;
  c->timing = false;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int st = e == hipSuccess ? LTB_OK : set_err(c, LTB_ERR_DEVICE, std::string("event: ") + hipGetErrorString(e));
  double best = 0;
  for (int r = -1; st == LTB_OK && r < reps; r++) {  // r = -1 warms up (and builds the light-cone table)
    float ms = 0;
    e = hipEventRecord(e0, c->stream);
    if (e == hipSuccess && (st = kpe_moments_impl(c, M, precision)) != LTB_OK) break;
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) st = set_err(c, LTB_ERR_DEVICE, std::string("recursion timing: ") + hipGetErrorString(e));
    if (r == 0 || (r > 0 && ms * 1e-3 < best)) best = ms * 1e-3;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  t.chunk_cols = saved_cols;
  c->timing = saved_timing;
  if (st == LTB_OK) *seconds = best;
  return st;
}

int kpe_resolved_tiling(ltb_ctx* c, int precision, ltb_kpe_tiling* out) {
  Plan p = precision == LTB_FP32_ACC64 ? make_plan<float>(c) : make_plan<double>(c);
  const int lo = c->tiling[precision == LTB_FP64 ? 1 : 0].lo_chunk_cols;
  // the width make_plan gives the low-order recursions (their chunk_cols = lo)
  out->lo_chunk_cols = lo > 0 ? (std::min(lo, c->norb) + p.W - 1) / p.W * p.W : p.ld;
  out->chunk_cols = p.ld;
  out->cols_per_lane = p.V;
  out->schedule = p.schedule;
  out->blocks_per_cu = p.schedule == LTB_SCHED_SWEEP ? p.slots * 8 / std::max(c->cu_count, 8) : 0;
  out->atoms_per_wave = p.schedule == LTB_SCHED_BLOCKED ? p.apw : 0;
  out->nontemporal = p.nt;
  out->order = p.schedule == LTB_SCHED_SWEEP ? p.order : 2;
  return LTB_OK;
}
