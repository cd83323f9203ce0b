// Neighbour list, H0 and Gamma assembled on the device from coordinates.
//
// Replaces the reference's host-side halo search and assembly
// (getHaloAtoms withSP2_deltaH.cpp:372-501, an O(N^2) scan over all atoms and
// images; ModifyNList :355-370; get4by4HmatrixElements :211-349; constructH
// :969-1066; initializeForSCF Gamma :854-900) by an O(N) cell list:
//   1. fractional coordinates and bins of width >= the 4 Bohr cutoff;
//   2. counting sort of the atoms by bin (atomics + scans);
//   3. per atom: count neighbours in the 27 surrounding bins (minimum image),
//      scan the counts into block-row pointers;
//   4. per atom: gather, sort by column, write the 4x4 Slater-Koster blocks
//      (diagonal block first), Gamma (-U_s = -1 on the diagonal, 1/r off it)
//      and the per-orbital diagonal / off-diagonal |H| sums for Gershgorin.
// The result is the pattern and values ltb_system_create builds on the host
// (same order, same formulas; device libm may differ in the last ulp).
// Periodic cells must be wider than twice the cutoff along every periodic
// row, so each pair has one image within range.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "quartz_model.hpp"

namespace {

constexpr double kRcut2 = 16.0;  // r^2 < 16 Bohr^2 (:392)
constexpr double kRmin2 = 0.1;   // r^2 > 0.1 (:421)
constexpr int kMaxNb = 20;       // maxN_A (:26)

// two-centre fit tables (:75-142): [0] Si-Si, [1] O-O, [2] Si-O
__constant__ double c_tab[3][20][4] = LTB_SK_TABLES;

struct Geom {
  double cell[9], inv[9];  // rows = periodic vectors; inv = inverse of that matrix
  double f0[3], span[3];
  int nbin[3];
  int mask;
};

// error codes raised by the kernels (first one wins)
enum { E_NONE = 0, E_CLOSE = 1, E_NEIGHBORS = 2 };

__device__ __forceinline__ int bin_of(const Geom& g, double f, int d) {
  double v = (f - g.f0[d]) / g.span[d];
  if (g.mask >> d & 1) v -= floor(v);
  int b = (int)floor(v * g.nbin[d]);
  return min(max(b, 0), g.nbin[d] - 1);
}

__global__ void k_frac(const double* __restrict__ xyz, int n, Geom g, double* __restrict__ f) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  for (int d = 0; d < 3; d++) f[3 * i + d] = x * g.inv[d] + y * g.inv[3 + d] + z * g.inv[6 + d];
}

// per-dimension min / max of the fractional coordinates (one block)
__global__ __launch_bounds__(256) void k_frac_range(const double* __restrict__ f, int n,
                                                    double* __restrict__ out) {
  __shared__ double lo[3][256], hi[3][256];
  double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < n; i += 256)
    for (int d = 0; d < 3; d++) {
      l[d] = fmin(l[d], f[3 * i + d]);
      h[d] = fmax(h[d], f[3 * i + d]);
    }
  for (int d = 0; d < 3; d++) {
    lo[d][threadIdx.x] = l[d];
    hi[d][threadIdx.x] = h[d];
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int d = 0; d < 3; d++) {
        lo[d][threadIdx.x] = fmin(lo[d][threadIdx.x], lo[d][threadIdx.x + w]);
        hi[d][threadIdx.x] = fmax(hi[d][threadIdx.x], hi[d][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 3) {
    out[threadIdx.x] = lo[threadIdx.x][0];
    out[3 + threadIdx.x] = hi[threadIdx.x][0];
  }
}

__global__ void k_bin_count(const double* __restrict__ f, int n, Geom g, int* __restrict__ bid,
                            int* __restrict__ bcount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = (bin_of(g, f[3 * i], 0) * g.nbin[1] + bin_of(g, f[3 * i + 1], 1)) * g.nbin[2] +
                bin_of(g, f[3 * i + 2], 2);
  bid[i] = b;
  atomicAdd(&bcount[b], 1);
}

__global__ void k_bin_fill(const int* __restrict__ bid, int n, int* __restrict__ cursor,
                           int* __restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  items[atomicAdd(&cursor[bid[i]], 1)] = i;
}

// Visit the atoms of the bins around atom i once each (a dimension with fewer
// than 3 bins is walked over all of its bins), calling fn(j, dv, r2) for
// every j != i with dv = r_j + image - r_i (minimum image on periodic rows).
template <typename F>
__device__ void for_neighbours(const Geom& g, const double* __restrict__ f, const int* __restrict__ head,
                               const int* __restrict__ items, int i, F fn) {
  const double fi[3] = {f[3 * i], f[3 * i + 1], f[3 * i + 2]};
  int bi[3], lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    bi[d] = bin_of(g, fi[d], d);
    if (g.nbin[d] < 3) {
      lo[d] = 0;
      hi[d] = g.nbin[d] - 1;
    } else {
      lo[d] = bi[d] - 1;
      hi[d] = bi[d] + 1;
    }
  }
  for (int b0 = lo[0]; b0 <= hi[0]; b0++)
    for (int b1 = lo[1]; b1 <= hi[1]; b1++)
      for (int b2 = lo[2]; b2 <= hi[2]; b2++) {
        int b[3] = {b0, b1, b2};
        bool ok = true;
        for (int d = 0; d < 3; d++) {
          if (g.nbin[d] < 3) continue;
          if (g.mask >> d & 1)
            b[d] = (b[d] + g.nbin[d]) % g.nbin[d];
          else if (b[d] < 0 || b[d] >= g.nbin[d])
            ok = false;
        }
        if (!ok) continue;
        const int bb = (b[0] * g.nbin[1] + b[1]) * g.nbin[2] + b[2];
        for (int k = head[bb]; k < head[bb + 1]; k++) {
          const int j = items[k];
          if (j == i) continue;
          double df[3];
          for (int d = 0; d < 3; d++) {
            df[d] = f[3 * j + d] - fi[d];
            if (g.mask >> d & 1) df[d] -= rint(df[d]);
          }
          double dv[3];
          for (int d = 0; d < 3; d++) dv[d] = df[0] * g.cell[d] + df[1] * g.cell[3 + d] + df[2] * g.cell[6 + d];
          fn(j, dv, dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        }
      }
}

__global__ void k_nb_count(const double* __restrict__ f, int n, Geom g, const int* __restrict__ head,
                           const int* __restrict__ items, int* __restrict__ rowlen, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cnt = 0, bad = E_NONE;
  for_neighbours(g, f, head, items, i, [&](int, const double*, double r2) {
    if (r2 <= kRmin2) bad = E_CLOSE;
    else if (r2 < kRcut2) cnt++;
  });
  if (bad == E_NONE && cnt > kMaxNb) bad = E_NEIGHBORS;
  if (bad != E_NONE) atomicCAS(err, E_NONE, bad * (1 << 28) + i);
  rowlen[i] = cnt + 1;  // + the diagonal block
}

// one 4x4 block of get4by4HmatrixElements (:211-349) for types zi, zj
// (14 Si / 8 O) and d = r_j - r_i; onsite constants when r < 0.01
__device__ void sk_block(int zi, int zj, const double d[3], double h[16]) {
  for (int k = 0; k < 16; k++) h[k] = 0;
  const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2], r = sqrt(r2);
  int t;
  if (zi != zj) {
    t = 2;
  } else if (zi == 14) {
    if (r < 0.01) {  // (:258-263)
      h[0] = LTB_ONSITE_SI_S;
      h[5] = h[10] = h[15] = LTB_ONSITE_SI_P;
      return;
    }
    t = 0;
  } else {
    if (r < 0.01) {  // (:299-304)
      h[0] = LTB_ONSITE_O_S;
      h[5] = h[10] = h[15] = LTB_ONSITE_O_P;
      return;
    }
    t = 1;
  }
  double S[4] = {0, 0, 0, 0};
  const double u = acos(r / 10);
  for (int p = 1; p <= 20; p++) {
    const double w = cos(p * u) * exp(-p * r);
    for (int c = 0; c < 4; c++) S[c] += c_tab[t][p - 1][c] * w;
  }
  h[0] = S[3];
  for (int a = 0; a < 3; a++) {
    h[1 + a] = S[2] * d[a] / r;
    h[4 * (1 + a)] = -S[2] * d[a] / r;
    for (int b = 0; b < 3; b++) {
      const double dd = d[a] * d[b] / r2;
      h[4 * (1 + a) + 1 + b] = S[0] * dd - S[1] * (dd - (a == b ? 1.0 : 0.0));
    }
  }
}

__global__ __launch_bounds__(128) void k_nb_fill(const double* __restrict__ f, int n, Geom g,
                                                 const int* __restrict__ head, const int* __restrict__ items,
                                                 const int* __restrict__ Z, const int* __restrict__ rowptr,
                                                 int* __restrict__ col, int* __restrict__ blkrow,
                                                 double* __restrict__ h0, double* __restrict__ gval,
                                                 double* __restrict__ diag0, double* __restrict__ offabs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int nj[kMaxNb];
  double nd[kMaxNb][3];
  int cnt = 0;
  for_neighbours(g, f, head, items, i, [&](int j, const double* dv, double r2) {
    if (r2 > kRmin2 && r2 < kRcut2 && cnt < kMaxNb) {
      nj[cnt] = j;
      nd[cnt][0] = dv[0];
      nd[cnt][1] = dv[1];
      nd[cnt][2] = dv[2];
      cnt++;
    }
  });
  // ascending columns (insertion sort, <= 20 entries)
  for (int a = 1; a < cnt; a++) {
    const int j = nj[a];
    const double d0 = nd[a][0], d1 = nd[a][1], d2 = nd[a][2];
    int b = a - 1;
    while (b >= 0 && nj[b] > j) {
      nj[b + 1] = nj[b];
      nd[b + 1][0] = nd[b][0];
      nd[b + 1][1] = nd[b][1];
      nd[b + 1][2] = nd[b][2];
      b--;
    }
    nj[b + 1] = j;
    nd[b + 1][0] = d0;
    nd[b + 1][1] = d1;
    nd[b + 1][2] = d2;
  }
  const int k0 = rowptr[i], zi = Z[i];
  double blk[16], off[4] = {0, 0, 0, 0}, dg[4];
  const double zero[3] = {0, 0, 0};
  sk_block(zi, zi, zero, blk);
  col[k0] = i;
  blkrow[k0] = i;
  gval[k0] = -1.0;  // -U_s, U_s = 1.0 (:876, :887)
  for (int e = 0; e < 16; e++) h0[16 * (size_t)k0 + e] = blk[e];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      if (a == b)
        dg[a] = blk[4 * a + b];
      else
        off[a] += fabs(blk[4 * a + b]);
    }
  for (int q = 0; q < cnt; q++) {
    const int k = k0 + 1 + q;
    sk_block(zi, Z[nj[q]], nd[q], blk);
    col[k] = nj[q];
    blkrow[k] = i;
    gval[k] = 1.0 / sqrt(nd[q][0] * nd[q][0] + nd[q][1] * nd[q][1] + nd[q][2] * nd[q][2]);  // (:890-896)
    for (int e = 0; e < 16; e++) h0[16 * (size_t)k + e] = blk[e];
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) off[a] += fabs(blk[4 * a + b]);
  }
  for (int a = 0; a < 4; a++) {
    diag0[4 * i + a] = dg[a];
    offabs[4 * i + a] = off[a];
  }
}

__global__ void k_iota(int* __restrict__ v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// exclusive scan of in[0..n) into out[0..n], out[n] = total: tiles of 1024
// scanned per block, tile sums scanned by one block, offsets added
__global__ __launch_bounds__(1024) void k_scan_tiles(const int* __restrict__ in, int n, int* __restrict__ out,
                                                     int* __restrict__ tile_sum) {
  __shared__ int sh[1024];
  const int i = blockIdx.x * 1024 + threadIdx.x;
  const int v = i < n ? in[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int t = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < n) out[i] = sh[threadIdx.x] - v;
  if (threadIdx.x == 1023) tile_sum[blockIdx.x] = sh[1023];
}

__global__ __launch_bounds__(1024) void k_scan_sums(int* __restrict__ tile_sum, int ntiles) {
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  __shared__ int sh[1024];
  for (int t0 = 0; t0 < ntiles; t0 += 1024) {
    const int t = t0 + threadIdx.x;
    const int v = t < ntiles ? tile_sum[t] : 0;
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
      const int x = threadIdx.x >= (unsigned)o ? sh[threadIdx.x - o] : 0;
      __syncthreads();
      sh[threadIdx.x] += x;
      __syncthreads();
    }
    if (t < ntiles) tile_sum[t] = carry + sh[threadIdx.x] - v;  // exclusive offset of tile t
    __syncthreads();
    if (threadIdx.x == 0) carry += sh[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) tile_sum[ntiles] = carry;
}

__global__ void k_scan_add(int* __restrict__ out, int n, const int* __restrict__ tile_sum, int ntiles) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += tile_sum[i / 1024];
  if (i == n) out[n] = tile_sum[ntiles];
}

bool invert3(const double a[9], double inv[9]) {
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
               c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(std::fabs(det) > 1e-12)) return false;
  const double id = 1.0 / det;
  inv[0] = c00 * id;
  inv[3] = c01 * id;
  inv[6] = c02 * id;
  inv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  inv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  inv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  inv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  inv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  inv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
  return true;
}

double width(const double c[9], int d) {
  const int e = (d + 1) % 3, f = (d + 2) % 3;
  const double cr[3] = {c[3 * e + 1] * c[3 * f + 2] - c[3 * e + 2] * c[3 * f + 1],
                        c[3 * e + 2] * c[3 * f + 0] - c[3 * e + 0] * c[3 * f + 2],
                        c[3 * e + 0] * c[3 * f + 1] - c[3 * e + 1] * c[3 * f + 0]};
  const double vol = std::fabs(c[3 * d] * cr[0] + c[3 * d + 1] * cr[1] + c[3 * d + 2] * cr[2]);
  return vol / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
}

int grid(int n, int b) { return (n + b - 1) / b; }

// exclusive scan on the stream; out has n + 1 entries
int scan(ltb_ctx* c, const int* in, int n, int* out, int* tiles) {
  const int nt = grid(n, 1024);
  hipLaunchKernelGGL(k_scan_tiles, dim3(nt), dim3(1024), 0, c->stream, in, n, out, tiles);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, c->stream, tiles, nt);
  hipLaunchKernelGGL(k_scan_add, dim3(grid(n + 1, 256)), dim3(256), 0, c->stream, out, n, tiles, nt);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

// scratch of one assembly, freed on every exit
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  hipError_t get(T** out, size_t count) {
    hipError_t e = hipMalloc((void**)out, sizeof(T) * std::max<size_t>(count, 1));
    if (e == hipSuccess) p.push_back((void*)*out);
    return e;
  }
};

}  // namespace

extern "C" {

int ltb_load_coords(ltb_ctx* c, int natoms, const int* Z, const double* xyz_bohr, const double* cell9,
                    int periodic_mask, double* seconds) {
  if (!c || natoms <= 0 || !Z || !xyz_bohr) return LTB_ERR_ARG;
  const int n = natoms;
  for (int i = 0; i < n; i++)
    if (Z[i] != 14 && Z[i] != 8) return set_err(c, LTB_ERR_ARG, "atomic number must be 14 (Si) or 8 (O)");
  for (size_t q = 0; q < 3 * (size_t)n; q++)
    if (!std::isfinite(xyz_bohr[q])) return set_err(c, LTB_ERR_NONFINITE, "non-finite coordinate");
  Geom g{};
  const double zero9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double* cell = cell9 ? cell9 : zero9;
  for (int k = 0; k < 9; k++) g.cell[k] = cell[k];
  g.mask = periodic_mask & 7;
  if (n >= (1 << 28)) return set_err(c, LTB_ERR_ARG, "too many atoms");
  const bool frac = invert3(g.cell, g.inv);
  if (!frac) {
    if (g.mask) return set_err(c, LTB_ERR_ARG, "singular periodic cell");
    // no periodicity: bin the Cartesian box (identity "cell")
    for (int k = 0; k < 9; k++) g.cell[k] = g.inv[k] = (k % 4 == 0) ? 1.0 : 0.0;
  }
  for (int d = 0; d < 3; d++)
    if ((g.mask >> d & 1) && width(g.cell, d) <= 2.0 * 4.0 + 1e-9)
      return set_err(c, LTB_ERR_ARG,
                     "periodic cell narrower than 8 Bohr: one image per pair is not guaranteed "
                     "(assemble on the host with ltb_system_create)");
  LTB_HIP(c, hipSetDevice(c->device));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  release_system(c);
  auto t0 = std::chrono::steady_clock::now();
  Scratch s;
  double *d_xyz, *d_f, *d_rng;
  int *d_Z, *d_bid, *d_err, *d_rowlen, *d_tiles;
  LTB_HIP(c, s.get(&d_xyz, 3 * (size_t)n));
  LTB_HIP(c, s.get(&d_f, 3 * (size_t)n));
  LTB_HIP(c, s.get(&d_rng, 6));
  LTB_HIP(c, s.get(&d_Z, n));
  LTB_HIP(c, s.get(&d_bid, n));
  LTB_HIP(c, s.get(&d_err, 1));
  LTB_HIP(c, s.get(&d_rowlen, n));
  LTB_HIP(c, hipMemcpyAsync(d_xyz, xyz_bohr, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  LTB_HIP(c, hipMemcpyAsync(d_Z, Z, sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_frac, dim3(grid(n, 256)), dim3(256), 0, c->stream, d_xyz, n, g, d_f);
  hipLaunchKernelGGL(k_frac_range, dim3(1), dim3(256), 0, c->stream, d_f, n, d_rng);
  double rng[6];
  LTB_HIP(c, hipMemcpyAsync(rng, d_rng, sizeof(rng), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  // bins of width >= the cutoff along each row (periodic rows span [0, 1))
  for (int d = 0; d < 3; d++) {
    const bool per = g.mask >> d & 1;
    g.f0[d] = per ? 0.0 : rng[d];
    g.span[d] = per ? 1.0 : std::max(rng[3 + d] - rng[d], 1e-9);
    const double w = frac ? width(g.cell, d) : 1.0;
    g.nbin[d] = std::min(1024, std::max(1, (int)std::floor(w * g.span[d] / 4.0)));
  }
  const size_t nbt = (size_t)g.nbin[0] * g.nbin[1] * g.nbin[2];
  if (nbt > (size_t)1 << 28) return set_err(c, LTB_ERR_ARG, "too many bins");
  int *d_bcount, *d_head, *d_cursor, *d_items;
  LTB_HIP(c, s.get(&d_bcount, nbt));
  LTB_HIP(c, s.get(&d_head, nbt + 1));
  LTB_HIP(c, s.get(&d_cursor, nbt + 1));
  LTB_HIP(c, s.get(&d_items, n));
  LTB_HIP(c, s.get(&d_tiles, grid((int)std::max<size_t>(nbt, n), 1024) + 1));
  LTB_HIP(c, hipMemsetAsync(d_bcount, 0, sizeof(int) * nbt, c->stream));
  LTB_HIP(c, hipMemsetAsync(d_err, 0, sizeof(int), c->stream));
  hipLaunchKernelGGL(k_bin_count, dim3(grid(n, 256)), dim3(256), 0, c->stream, d_f, n, g, d_bid, d_bcount);
  int st = scan(c, d_bcount, (int)nbt, d_head, d_tiles);
  if (st) return st;
  LTB_HIP(c, hipMemcpyAsync(d_cursor, d_head, sizeof(int) * nbt, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_bin_fill, dim3(grid(n, 256)), dim3(256), 0, c->stream, d_bid, n, d_cursor, d_items);
  hipLaunchKernelGGL(k_nb_count, dim3(grid(n, 128)), dim3(128), 0, c->stream, d_f, n, g, d_head, d_items,
                     d_rowlen, d_err);
  LTB_HIP(c, hipGetLastError());
  int err = 0;
  LTB_HIP(c, hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  if (err) {
    const int code = err >> 28, atom = err & ((1 << 28) - 1);
    if (code == E_CLOSE)
      return set_err(c, LTB_ERR_GEOMETRY, "atom " + std::to_string(atom) + " has a neighbour closer than sqrt(0.1) Bohr");
    return set_err(c, LTB_ERR_NEIGHBORS, "atom " + std::to_string(atom) + " has more than 20 neighbours");
  }
  // block-row pointers, then the system buffers
  int* d_rowptr_tmp;
  LTB_HIP(c, s.get(&d_rowptr_tmp, n + 1));
  if ((st = scan(c, d_rowlen, n, d_rowptr_tmp, d_tiles))) return st;
  int nnzb = 0;
  LTB_HIP(c, hipMemcpyAsync(&nnzb, d_rowptr_tmp + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  if ((st = alloc_system(c, n, nnzb, nnzb))) {
    release_system(c);
    return st;
  }
  LTB_HIP(c, hipMemcpyAsync(c->d_rowptr, d_rowptr_tmp, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, c->stream));
  LTB_HIP(c, hipMemcpyAsync(c->d_grow, d_rowptr_tmp, sizeof(int) * (n + 1), hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_nb_fill, dim3(grid(n, 128)), dim3(128), 0, c->stream, d_f, n, g, d_head, d_items, d_Z,
                     c->d_rowptr, c->d_col, c->d_blkrow, c->d_h0, c->d_gval, c->d_diag0, c->d_offabs);
  LTB_HIP(c, hipGetLastError());
  LTB_HIP(c, hipMemcpyAsync(c->d_gcol, c->d_col, sizeof(int) * nnzb, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_iota, dim3(grid(n, 256)), dim3(256), 0, c->stream, c->d_visit, n);
  // neutral per-spin populations: Si 2, O 3
  std::vector<double> p0(n);
  double ne = 0;
  for (int i = 0; i < n; i++) {
    p0[i] = Z[i] == 14 ? 2.0 : 3.0;
    ne += 2.0 * p0[i];
  }
  LTB_HIP(c, hipMemcpyAsync(c->d_p0, p0.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  LTB_HIP(c, hipMemcpyAsync(c->d_pin, p0.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  c->nat = n;
  c->norb = 4 * n;
  c->nnzb = nnzb;
  c->gnnz = nnzb;
  c->nelectrons = ne;
  if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return LTB_OK;
}

int ltb_get_sizes(ltb_ctx* c, int* natoms, int* nnzb, int* gamma_nnz, double* nelectrons) {
  if (!c) return LTB_ERR_ARG;
  if (natoms) *natoms = c->nat;
  if (nnzb) *nnzb = c->nnzb;
  if (gamma_nnz) *gamma_nnz = c->gnnz;
  if (nelectrons) *nelectrons = c->nelectrons;
  return LTB_OK;
}

int ltb_get_hamiltonian(ltb_ctx* c, int* rowptr, int* col, double* val16) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  if (rowptr) LTB_HIP(c, hipMemcpy(rowptr, c->d_rowptr, sizeof(int) * (c->nat + 1), hipMemcpyDeviceToHost));
  if (col) LTB_HIP(c, hipMemcpy(col, c->d_col, sizeof(int) * c->nnzb, hipMemcpyDeviceToHost));
  if (val16) LTB_HIP(c, hipMemcpy(val16, c->d_h0, sizeof(double) * 16 * c->nnzb, hipMemcpyDeviceToHost));
  return LTB_OK;
}

int ltb_get_gamma(ltb_ctx* c, int* rowptr, int* col, double* val) {
  if (!c) return LTB_ERR_ARG;
  if (c->nat <= 0) return set_err(c, LTB_ERR_STATE, "no system loaded");
  if (rowptr) LTB_HIP(c, hipMemcpy(rowptr, c->d_grow, sizeof(int) * (c->nat + 1), hipMemcpyDeviceToHost));
  if (col) LTB_HIP(c, hipMemcpy(col, c->d_gcol, sizeof(int) * c->gnnz, hipMemcpyDeviceToHost));
  if (val) LTB_HIP(c, hipMemcpy(val, c->d_gval, sizeof(double) * c->gnnz, hipMemcpyDeviceToHost));
  return LTB_OK;
}

}  // extern "C"
