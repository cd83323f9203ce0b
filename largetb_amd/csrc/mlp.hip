// Learned SCC starting guess: per-atom descriptors and a per-species MLP,
// trained on the device with hand-written forward, backward and Adam kernels.
// No reference counterpart (the reference always starts from the neutral
// charges of arrangeCoordinates, withSP2_deltaH.cpp:714-720).
//
// Descriptor (80 features per atom, Bohr units, cutoff Rc = 6 A):
//   fc(r) = (cos(pi r / Rc) + 1) / 2
//   radial [32 t + k], neighbour species t (0 Si, 1 O), k = 0..31:
//     sum_{j in t} exp(-eta (r_ij - mu_k)^2) fc(r_ij),
//     mu_k = r0 + k (Rc - r0)/31, r0 = 0.5 A, eta = 1 / (2 ((Rc - r0)/31)^2)
//   angular [64 + 4 z + 2 l + e], zeta = {1,2,4,8}[z], lambda = {-1,+1}[l],
//     eta_a = {0.01, 0.04}[e] Bohr^-2, over unordered neighbour pairs j < k:
//     2^(1-zeta) (1 + lambda cos theta_jik)^zeta exp(-eta_a (r_ij^2 + r_ik^2)) fc_ij fc_ik
// Network, one per centre species: x^ = (x - mean_c) / sd_c, three SiLU layers
// of 64, linear output y; q_raw = shift_c + y; neutrality by mean projection
// over each structure: q = q_raw - (sum q_raw - Q) / n.
// Parameter vector (floats): per species [W1 64x80 | b1 | W2 64x64 | b2 |
// W3 64x64 | b3 | w4 64 | b4] (13569 each, trainable), then mean[2][80],
// sd[2][80], shift[2] (fixed after fit_normalization).
//
// Evaluation on MFMA: the atoms of one species form tiles of 64; a workgroup
// (4 waves) stages its species' weights and one tile's normalised features in
// LDS and forms every 64-wide layer as a 64 x 64 product on
// v_mfma_f32_32x32x2f32, one 32 x 32 quadrant per wave.  The backward pass
// recomputes the forward pass of its tile (no per-atom activations are kept
// in HBM), overwrites z3, z2, z1 in LDS with the layer gradients and adds
// the tile's weight gradients to MFMA accumulators that stay in registers
// over all the tiles of a persistent workgroup; every workgroup then writes
// one partial gradient and a fixed-order sum forms the gradient (the result
// does not depend on timing).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ctx.hpp"
#include "system.hpp"

namespace {

constexpr int NF = 80, NH = 64;
constexpr int OW1 = 0, OB1 = OW1 + NH * NF, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NH, PS = OB4 + 1;  // 13569
constexpr int NTRAIN = 2 * PS;
constexpr int OMEAN = NTRAIN, OSD = OMEAN + 2 * NF, OSHIFT = OSD + 2 * NF, NPARAM = OSHIFT + 2;
constexpr double kPi = 3.141592653589793238463;
constexpr double kRcAng = 6.0, kR0Ang = 0.5;
constexpr int NPB = 256;   // persistent workgroups of the backward pass (partial gradients)
constexpr int TA = 64;     // atoms per tile
constexpr int LX = NF + 1, LH = NH + 1;  // odd LDS row strides: row and column reads are conflict-free

// the logistic function through the hardware reciprocal (about 1 ulp)
__device__ __forceinline__ float sigm(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigm(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ int species(int Z) { return Z == 14 ? 0 : 1; }

__global__ __launch_bounds__(128) void k_descriptors(int nat, const int* __restrict__ Z, const int* __restrict__ rowptr,
                              const int* __restrict__ nbj, const double* __restrict__ nbd,
                              double rc, double r0, float* __restrict__ feat) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nat) return;
  float G[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) G[f] = 0.f;
  const double dmu = (rc - r0) / 31.0;
  const float eta = (float)(1.0 / (2.0 * dmu * dmu));
  const int pb = rowptr[i], pe = rowptr[i + 1];
  for (int p = pb; p < pe; p++) {
    const double dx = nbd[3 * p], dy = nbd[3 * p + 1], dz = nbd[3 * p + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    if (r >= rc) continue;
    const float fc = (float)(0.5 * (cos(kPi * r / rc) + 1.0));
    const int t = species(Z[nbj[p]]);
#pragma unroll
    for (int k = 0; k < 32; k++) {
      const float d = (float)(r - (r0 + k * dmu));
      const float g = __expf(-eta * d * d) * fc;
      if (t == 0)
        G[k] += g;
      else
        G[32 + k] += g;
    }
  }
  const float etaa[2] = {0.01f, 0.04f};
  for (int p = pb; p < pe; p++) {
    const double ax = nbd[3 * p], ay = nbd[3 * p + 1], az = nbd[3 * p + 2];
    const double ra = sqrt(ax * ax + ay * ay + az * az);
    if (ra >= rc) continue;
    const float fa = (float)(0.5 * (cos(kPi * ra / rc) + 1.0));
    for (int q = p + 1; q < pe; q++) {
      const double bx = nbd[3 * q], by = nbd[3 * q + 1], bz = nbd[3 * q + 2];
      const double rb = sqrt(bx * bx + by * by + bz * bz);
      if (rb >= rc) continue;
      const float fb = (float)(0.5 * (cos(kPi * rb / rc) + 1.0));
      const float ct = (float)((ax * bx + ay * by + az * bz) / (ra * rb));
      const float r2 = (float)(ra * ra + rb * rb);
#pragma unroll
      for (int l = 0; l < 2; l++) {
        const float base = 1.0f + (l == 0 ? -ct : ct);
        float pw = base;  // zeta = 1
#pragma unroll
        for (int z = 0; z < 4; z++) {
          if (z > 0) pw = pw * pw;  // zeta = 2, 4, 8
          const float pref = (z == 0) ? 1.0f : (z == 1 ? 0.5f : (z == 2 ? 0.125f : 0.0078125f));
#pragma unroll
          for (int e = 0; e < 2; e++) G[64 + 4 * z + 2 * l + e] += pref * pw * __expf(-etaa[e] * r2) * fa * fb;
        }
      }
    }
  }
#pragma unroll
  for (int f = 0; f < NF; f++) feat[(size_t)i * NF + f] = G[f];
}

typedef float acc16 __attribute__((ext_vector_type(16)));
__device__ __forceinline__ acc16 mma32(float a, float b, acc16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}
// v_mfma_f32_32x32x2f32 lane map: A[m = lane % 32][k = lane / 32],
// B[k = lane / 32][n = lane % 32]; accumulator e of a lane holds row
// 8 (e / 4) + 4 (lane / 32) + e % 4 and column lane % 32 of the 32 x 32 tile.
__device__ __forceinline__ int acc_row(int lane, int e) { return 8 * (e >> 2) + 4 * (lane >> 5) + (e & 3); }

struct MlpSmem {
  float W1[NH * LX];  // W1[o][f] at o LX + f
  float W2[NH * LH];
  float W3[NH * LH];
  float b[3][NH];
  float w4[NH];
  float mean[NF], sd[NF];
  float X[TA * LX];     // normalised features of the tile's atoms (zero rows past the species' end)
  float Z[3][TA * LH];  // z1, z2, z3 of the tile; the backward pass overwrites them with G1, G2, G3
  float g[TA];          // dL/dq_raw of the tile's atoms (zero past the end)
};

__device__ void stage_weights(MlpSmem& sm, const float* __restrict__ P, int c) {
  const float* W = P + c * PS;
  for (int e = threadIdx.x; e < NH * NF; e += 256) sm.W1[(e / NF) * LX + e % NF] = W[OW1 + e];
  for (int e = threadIdx.x; e < NH * NH; e += 256) {
    sm.W2[(e / NH) * LH + e % NH] = W[OW2 + e];
    sm.W3[(e / NH) * LH + e % NH] = W[OW3 + e];
  }
  for (int e = threadIdx.x; e < NH; e += 256) {
    sm.b[0][e] = W[OB1 + e];
    sm.b[1][e] = W[OB2 + e];
    sm.b[2][e] = W[OB3 + e];
    sm.w4[e] = W[OW4 + e];
  }
  for (int e = threadIdx.x; e < NF; e += 256) {
    sm.mean[e] = P[OMEAN + c * NF + e];
    sm.sd[e] = P[OSD + c * NF + e];
  }
  __syncthreads();  // the tile loads read mean and sd
}

// quadrant (wr, wc) of in[TA x K] W[NH x K]^T (silu applied to `in` on load)
template <int K, bool SILU_IN>
__device__ __forceinline__ acc16 layer_quadrant(const float* in, int li, const float* W, int lw, int wr, int wc,
                                                int lane) {
  acc16 acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.f;
  const int m = wr * 32 + (lane & 31), n = wc * 32 + (lane & 31), kq = lane >> 5;
#pragma unroll 8
  for (int k0 = 0; k0 < K; k0 += 2) {
    float a = in[m * li + k0 + kq];
    if (SILU_IN) a = silu(a);
    acc = mma32(a, W[n * lw + k0 + kq], acc);
  }
  return acc;
}

__device__ __forceinline__ void store_quadrant(float* out, const float* bias, acc16 acc, int wr, int wc, int lane) {
  const int col = wc * 32 + (lane & 31);
#pragma unroll
  for (int e = 0; e < 16; e++) out[(wr * 32 + acc_row(lane, e)) * LH + col] = acc[e] + bias[col];
}

// one tile's raw features (and g) in registers: thread i holds elements
// i, i + 256, ... of the tile's 64 x 80 block, so the next tile's gathers are in
// flight while the current tile computes
constexpr int NXR = TA * NF / 256;
static_assert(TA * NF % 256 == 0, "whole feature block per thread");
struct TileRegs {
  float x[NXR];
  float g;
};

template <bool BWD>
__device__ __forceinline__ void tile_fetch(TileRegs& tr, const int* __restrict__ idx, int nidx, int t, int ntile,
                                           const float* __restrict__ feat, const double* __restrict__ gq) {
#pragma unroll
  for (int j = 0; j < NXR; j++) {
    const int e = threadIdx.x + 256 * j, r = e / NF, f = e - r * NF, k = t * TA + r;
    tr.x[j] = (t < ntile && k < nidx) ? feat[(size_t)idx[k] * NF + f] : 0.f;
  }
  if (BWD) {
    const int k = t * TA + threadIdx.x;
    tr.g = (threadIdx.x < TA && t < ntile && k < nidx) ? (float)gq[idx[k]] : 0.f;
  }
}

// the tile's normalised features (and g) into LDS, then z1, z2, z3
template <bool BWD>
__device__ void tile_forward(MlpSmem& sm, const TileRegs& tr, int nidx, int t, int wave, int lane) {
  const int wr = wave >> 1, wc = wave & 1;
#pragma unroll
  for (int j = 0; j < NXR; j++) {
    const int e = threadIdx.x + 256 * j, r = e / NF, f = e - r * NF;
    sm.X[r * LX + f] = t * TA + r < nidx ? (tr.x[j] - sm.mean[f]) / sm.sd[f] : 0.f;
  }
  if (BWD && threadIdx.x < TA) sm.g[threadIdx.x] = tr.g;
  __syncthreads();
  store_quadrant(sm.Z[0], sm.b[0], layer_quadrant<NF, false>(sm.X, LX, sm.W1, LX, wr, wc, lane), wr, wc, lane);
  __syncthreads();
  store_quadrant(sm.Z[1], sm.b[1], layer_quadrant<NH, true>(sm.Z[0], LH, sm.W2, LH, wr, wc, lane), wr, wc, lane);
  __syncthreads();
  store_quadrant(sm.Z[2], sm.b[2], layer_quadrant<NH, true>(sm.Z[1], LH, sm.W3, LH, wr, wc, lane), wr, wc, lane);
  __syncthreads();
}

// sum over the 64 lanes of a wave (fixed order; every lane gets the sum)
__device__ __forceinline__ float wave_sum(float v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// column sums over the tile's 64 rows of M (LDS, row stride LH) for the 16
// columns 16 wave .. 16 wave + 15: lane l adds rows 16 (l / 16) .. + 15 of column
// 16 wave + l % 16, then the four row groups are added across lanes; every lane
// returns the sum of its column.  G_WEIGHT: sum_r g_r silu(M[r][o]) instead.
template <bool G_WEIGHT>
__device__ __forceinline__ float colsum16(const MlpSmem& sm, const float* M, int wave, int lane) {
  const int col = 16 * wave + (lane & 15), r0 = 16 * (lane >> 4);
  float s = 0.f;
#pragma unroll
  for (int r = r0; r < r0 + 16; r++) s = G_WEIGHT ? fmaf(sm.g[r], silu(M[r * LH + col]), s) : s + M[r * LH + col];
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

// q_raw of every atom of species c (idx: its atoms, nidx of them)
__global__ __launch_bounds__(256) void k_mlp_forward(int c, const int* __restrict__ idx, int nidx,
                                                     const float* __restrict__ feat, const float* __restrict__ P,
                                                     double* __restrict__ qraw) {
  __shared__ MlpSmem sm;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ntile = (nidx + TA - 1) / TA;
  if ((int)blockIdx.x >= ntile) return;
  TileRegs tr;
  tile_fetch<false>(tr, idx, nidx, blockIdx.x, ntile, feat, nullptr);
  stage_weights(sm, P, c);
  const float b4 = P[c * PS + OB4], shift = P[OSHIFT + c];
  for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
    tile_forward<false>(sm, tr, nidx, t, wave, lane);
    tile_fetch<false>(tr, idx, nidx, t + gridDim.x, ntile, feat, nullptr);  // the next tile
    if (threadIdx.x < TA && t * TA + (int)threadIdx.x < nidx) {
      const float* z3 = sm.Z[2] + threadIdx.x * LH;
      float y = b4;
#pragma unroll 8
      for (int f = 0; f < NH; f++) y = fmaf(sm.w4[f], silu(z3[f]), y);
      qraw[idx[t * TA + threadIdx.x]] = (double)shift + (double)y;
    }
    __syncthreads();  // the next tile overwrites X and Z
  }
}

// dL/dW of species c over its atoms: workgroup p adds the tiles p, p + NPB, ...
// and writes its sums to partial[p][c PS ...] (zeros when it has no tile)
__global__ __launch_bounds__(256) void k_mlp_backward(int c, const int* __restrict__ idx, int nidx,
                                                      const float* __restrict__ feat, const float* __restrict__ P,
                                                      const double* __restrict__ gq, float* __restrict__ partial) {
  __shared__ MlpSmem sm;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int wr = wave >> 1, wc = wave & 1, tid = threadIdx.x;
  const int ntile = (nidx + TA - 1) / TA;
  TileRegs tr;
  tile_fetch<true>(tr, idx, nidx, blockIdx.x, ntile, feat, gq);
  stage_weights(sm, P, c);
  acc16 dw3, dw2, dw1a, dw1b;  // dW3, dW2 quadrant (wr, wc); dW1 tiles q = wave and q = wave + 4 (waves 0, 1)
#pragma unroll
  for (int e = 0; e < 16; e++) dw3[e] = dw2[e] = dw1a[e] = dw1b[e] = 0.f;
  // bias gradients: lane l of wave w holds those of output 16 w + l % 16
  float db1 = 0.f, db2 = 0.f, db3 = 0.f, dw4 = 0.f, db4 = 0.f;
  const int m32 = lane & 31, kq = lane >> 5;
  // dW1 tile q: rows (outputs) 32 (q / 3), columns (features) 32 (q % 3), features >= NF are zero
  const int qa = wave, qb = wave + 4;
  for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
    tile_forward<true>(sm, tr, nidx, t, wave, lane);
    tile_fetch<true>(tr, idx, nidx, t + gridDim.x, ntile, feat, gq);  // the next tile
    // layer 4: dw4 = sum_r g_r silu(z3), db4 = sum_r g_r
    dw4 += colsum16<true>(sm, sm.Z[2], wave, lane);
    if (wave == 0) db4 += wave_sum(sm.g[lane]);
    __syncthreads();
    // G3 = g w4 silu'(z3), in place of z3
    for (int e = tid; e < TA * NH; e += 256) {
      const int r = e / NH, o = e - r * NH;
      float* z = &sm.Z[2][r * LH + o];
      *z = sm.g[r] * sm.w4[o] * dsilu(*z);
    }
    __syncthreads();
    // dW3 += G3^T silu(z2); P2 = G3 W3; db3 += column sums of G3
    {
      acc16 p2;
#pragma unroll
      for (int e = 0; e < 16; e++) p2[e] = 0.f;
#pragma unroll 8
      for (int k0 = 0; k0 < TA; k0 += 2) {
        const int r = k0 + kq;
        dw3 = mma32(sm.Z[2][r * LH + wr * 32 + m32], silu(sm.Z[1][r * LH + wc * 32 + m32]), dw3);
      }
#pragma unroll 8
      for (int k0 = 0; k0 < NH; k0 += 2) {
        const int o = k0 + kq;
        p2 = mma32(sm.Z[2][(wr * 32 + m32) * LH + o], sm.W3[o * LH + wc * 32 + m32], p2);
      }
      db3 += colsum16<false>(sm, sm.Z[2], wave, lane);
      __syncthreads();
      // G2 = P2 silu'(z2), in place of z2 (each lane its own elements)
      const int col = wc * 32 + m32;
#pragma unroll
      for (int e = 0; e < 16; e++) {
        float* z = &sm.Z[1][(wr * 32 + acc_row(lane, e)) * LH + col];
        *z = p2[e] * dsilu(*z);
      }
      __syncthreads();
    }
    // dW2 += G2^T silu(z1); P1 = G2 W2; db2 += column sums of G2
    {
      acc16 p1;
#pragma unroll
      for (int e = 0; e < 16; e++) p1[e] = 0.f;
#pragma unroll 8
      for (int k0 = 0; k0 < TA; k0 += 2) {
        const int r = k0 + kq;
        dw2 = mma32(sm.Z[1][r * LH + wr * 32 + m32], silu(sm.Z[0][r * LH + wc * 32 + m32]), dw2);
      }
#pragma unroll 8
      for (int k0 = 0; k0 < NH; k0 += 2) {
        const int o = k0 + kq;
        p1 = mma32(sm.Z[1][(wr * 32 + m32) * LH + o], sm.W2[o * LH + wc * 32 + m32], p1);
      }
      db2 += colsum16<false>(sm, sm.Z[1], wave, lane);
      __syncthreads();
      const int col = wc * 32 + m32;
#pragma unroll
      for (int e = 0; e < 16; e++) {
        float* z = &sm.Z[0][(wr * 32 + acc_row(lane, e)) * LH + col];
        *z = p1[e] * dsilu(*z);
      }
      __syncthreads();
    }
    // dW1 += G1^T X; db1 += column sums of G1
#pragma unroll 8
    for (int k0 = 0; k0 < TA; k0 += 2) {
      const int r = k0 + kq;
      const float* g1 = sm.Z[0] + r * LH;
      const float* x = sm.X + r * LX;
      const int fa = 32 * (qa % 3) + m32;
      dw1a = mma32(g1[32 * (qa / 3) + m32], fa < NF ? x[fa] : 0.f, dw1a);
      if (wave < 2) {
        const int fb = 32 * (qb % 3) + m32;
        dw1b = mma32(g1[32 * (qb / 3) + m32], fb < NF ? x[fb] : 0.f, dw1b);
      }
    }
    db1 += colsum16<false>(sm, sm.Z[0], wave, lane);
    __syncthreads();  // the next tile overwrites X, Z and g
  }
  float* out = partial + (size_t)blockIdx.x * NTRAIN + c * PS;
  {
    const int col = wc * 32 + m32;
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const int row = wr * 32 + acc_row(lane, e);
      out[OW3 + row * NH + col] = dw3[e];
      out[OW2 + row * NH + col] = dw2[e];
    }
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const int f = 32 * (qa % 3) + m32;
      if (f < NF) out[OW1 + (32 * (qa / 3) + acc_row(lane, e)) * NF + f] = dw1a[e];
    }
    if (wave < 2) {
#pragma unroll
      for (int e = 0; e < 16; e++) {
        const int f = 32 * (qb % 3) + m32;
        if (f < NF) out[OW1 + (32 * (qb / 3) + acc_row(lane, e)) * NF + f] = dw1b[e];
      }
    }
  }
  if (lane < 16) {
    const int o = 16 * wave + lane;
    out[OB1 + o] = db1;
    out[OB2 + o] = db2;
    out[OB3 + o] = db3;
    out[OW4 + o] = dw4;
  }
  if (tid == 0) out[OB4] = db4;
}

// mean projection per structure: q = q_raw - (sum q_raw - Q) / n_seg
__global__ __launch_bounds__(256) void k_project(const int* __restrict__ seg,
                                                 const double* __restrict__ qtot,
                                                 const double* __restrict__ qraw,
                                                 double* __restrict__ q) {
  const int s = blockIdx.x, b = seg[s], e = seg[s + 1];
  __shared__ double sh[256];
  double acc = 0;
  for (int a = b + threadIdx.x; a < e; a += 256) acc += qraw[a];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double shift = (sh[0] - qtot[s]) / (double)(e - b);
  for (int a = b + threadIdx.x; a < e; a += 256) q[a] = qraw[a] - shift;
}

// loss = sum (q - t)^2 / n_total; g = dloss/dq_raw (through the projection)
__global__ __launch_bounds__(256) void k_loss_grad(const int* __restrict__ seg,
                                                   const double* __restrict__ q,
                                                   const double* __restrict__ t, int ntot,
                                                   double* __restrict__ g,
                                                   double* __restrict__ loss_part) {
  const int s = blockIdx.x, b = seg[s], e = seg[s + 1];
  __shared__ double sh[256], sl[256];
  double acc = 0, lacc = 0;
  for (int a = b + threadIdx.x; a < e; a += 256) {
    const double d = q[a] - t[a];
    acc += 2.0 * d / ntot;
    lacc += d * d / ntot;
  }
  sh[threadIdx.x] = acc;
  sl[threadIdx.x] = lacc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x] += sh[threadIdx.x + w];
      sl[threadIdx.x] += sl[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double mean = sh[0] / (double)(e - b);
  for (int a = b + threadIdx.x; a < e; a += 256) g[a] = 2.0 * (q[a] - t[a]) / ntot - mean;
  if (threadIdx.x == 0) loss_part[s] = sl[0];
}

// *out = sum of v[0..n) (one workgroup of 256, fixed order)
__global__ __launch_bounds__(256) void k_sum(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) s += v[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__global__ void k_reduce_partials(const float* __restrict__ partial, float* __restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NTRAIN) return;
  float s = 0;
  for (int p = 0; p < NPB; p++) s += partial[(size_t)p * NTRAIN + j];
  grad[j] = s;
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NTRAIN) return;
  const float gj = g[j];
  const float mj = b1 * m[j] + (1 - b1) * gj;
  const float vj = b2 * v[j] + (1 - b2) * gj * gj;
  m[j] = mj;
  v[j] = vj;
  P[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

}  // namespace

struct ltb_mlp {
  ltb_ctx* ctx = nullptr;
  float *d_P = nullptr, *d_grad = nullptr, *d_m = nullptr, *d_v = nullptr, *d_partial = nullptr;
  int adam_t = 0;
  int n = 0, nseg = 0;
  float* d_feat = nullptr;
  int *d_Z = nullptr, *d_seg = nullptr, *d_idx[2] = {nullptr, nullptr};
  int nidx[2] = {0, 0};
  double *d_qtot = nullptr, *d_t = nullptr, *d_qraw = nullptr, *d_q = nullptr, *d_g = nullptr,
         *d_lpart = nullptr, *d_loss = nullptr;
  bool has_target = false;
};

namespace {

void free_data(ltb_mlp* m) {
  void* ps[] = {m->d_feat, m->d_Z,
                m->d_seg, m->d_idx[0], m->d_idx[1], m->d_qtot, m->d_t, m->d_qraw, m->d_q, m->d_g,
                m->d_lpart};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->d_feat = nullptr;
  m->d_Z = m->d_seg = m->d_idx[0] = m->d_idx[1] = nullptr;
  m->d_qtot = m->d_t = m->d_qraw = m->d_q = m->d_g = m->d_lpart = nullptr;
  m->n = m->nseg = 0;
}

int upload_data(ltb_mlp* m, int n, const float* feat, const int* Z, int nseg, const int* seg, const double* qtot,
                const double* target, const std::vector<int> (&idx)[2]);

void free_params(ltb_mlp* m) {
  void* ps[] = {m->d_P, m->d_grad, m->d_m, m->d_v, m->d_partial, m->d_loss};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->d_P = m->d_grad = m->d_m = m->d_v = m->d_partial = nullptr;
  m->d_loss = nullptr;
}

int descriptors_device(ltb_ctx* c, const ltb_system* s, float* d_feat) {
  std::vector<int> rp, nj;
  std::vector<double> nd;
  ltb::Status st = ltb::pairs_within(*s, kRcAng * ltb::kAngToBohr, rp, nj, nd);
  if (st.code) return set_err(c, st.code, st.msg);
  int *d_rp = nullptr, *d_nj = nullptr, *d_Z = nullptr;
  double* d_nd = nullptr;
  const int nat = s->natoms;
  LTB_HIP(c, hipMalloc(&d_rp, sizeof(int) * (nat + 1)));
  LTB_HIP(c, hipMalloc(&d_nj, sizeof(int) * std::max<size_t>(nj.size(), 1)));
  LTB_HIP(c, hipMalloc(&d_nd, sizeof(double) * std::max<size_t>(nd.size(), 1)));
  LTB_HIP(c, hipMalloc(&d_Z, sizeof(int) * nat));
  LTB_HIP(c, hipMemcpyAsync(d_rp, rp.data(), sizeof(int) * (nat + 1), hipMemcpyHostToDevice, c->stream));
  if (!nj.empty()) {
    LTB_HIP(c, hipMemcpyAsync(d_nj, nj.data(), sizeof(int) * nj.size(), hipMemcpyHostToDevice, c->stream));
    LTB_HIP(c, hipMemcpyAsync(d_nd, nd.data(), sizeof(double) * nd.size(), hipMemcpyHostToDevice, c->stream));
  }
  LTB_HIP(c, hipMemcpyAsync(d_Z, s->Z.data(), sizeof(int) * nat, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_descriptors, dim3((nat + 127) / 128), dim3(128), 0, c->stream, nat, d_Z, d_rp,
                     d_nj, d_nd, kRcAng * ltb::kAngToBohr, kR0Ang * ltb::kAngToBohr, d_feat);
  LTB_HIP(c, hipGetLastError());
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  (void)hipFree(d_rp);
  (void)hipFree(d_nj);
  (void)hipFree(d_nd);
  (void)hipFree(d_Z);
  return LTB_OK;
}

// forward tiles: up to 4 workgroups per CU's worth of tiles in flight
int forward_grid(const ltb_mlp* m, int nidx) {
  const int ntile = (nidx + TA - 1) / TA;
  return std::max(1, std::min(ntile, 4 * std::max(m->ctx->cu_count, 1)));
}

int forward_async(ltb_mlp* m) {
  ltb_ctx* c = m->ctx;
  for (int sp = 0; sp < 2; sp++) {
    if (m->nidx[sp] == 0) continue;
    hipLaunchKernelGGL(k_mlp_forward, dim3(forward_grid(m, m->nidx[sp])), dim3(256), 0, c->stream, sp,
                       m->d_idx[sp], m->nidx[sp], m->d_feat, m->d_P, m->d_qraw);
  }
  hipLaunchKernelGGL(k_project, dim3(m->nseg), dim3(256), 0, c->stream, m->d_seg, m->d_qtot, m->d_qraw,
                     m->d_q);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int loss_grad_async(ltb_mlp* m) {
  ltb_ctx* c = m->ctx;
  hipLaunchKernelGGL(k_loss_grad, dim3(m->nseg), dim3(256), 0, c->stream, m->d_seg, m->d_q, m->d_t,
                     m->n, m->d_g, m->d_lpart);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, c->stream, m->d_lpart, m->nseg, m->d_loss);
  // both species always run: a species without atoms writes zero partials
  for (int sp = 0; sp < 2; sp++)
    hipLaunchKernelGGL(k_mlp_backward, dim3(NPB), dim3(256), 0, c->stream, sp, m->d_idx[sp], m->nidx[sp],
                       m->d_feat, m->d_P, m->d_g, m->d_partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3((NTRAIN + 255) / 256), dim3(256), 0, c->stream,
                     m->d_partial, m->d_grad);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

}  // namespace

extern "C" {

int ltb_mlp_num_params(void) { return NPARAM; }

int ltb_mlp_create(ltb_ctx* c, unsigned long long seed, ltb_mlp** out) {
  if (!c || !out) return LTB_ERR_ARG;
  auto* m = new ltb_mlp();
  m->ctx = c;
  std::vector<float> P(NPARAM, 0.f);
  uint64_t st = seed;
  auto uni = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * 0x1.0p-53 * 2.0 - 1.0;  // (-1, 1)
  };
  for (int sp = 0; sp < 2; sp++) {
    float* W = P.data() + sp * PS;
    const double a1 = std::sqrt(6.0 / (NF + NH)), a2 = std::sqrt(6.0 / (NH + NH));
    for (int k = 0; k < NH * NF; k++) W[OW1 + k] = (float)(a1 * uni());
    for (int k = 0; k < NH * NH; k++) W[OW2 + k] = (float)(a2 * uni());
    for (int k = 0; k < NH * NH; k++) W[OW3 + k] = (float)(a2 * uni());
    for (int k = 0; k < NH; k++) W[OW4 + k] = (float)(0.1 * a2 * uni());
  }
  for (int k = 0; k < 2 * NF; k++) P[OSD + k] = 1.f;
  P[OSHIFT] = 4.0f;      // neutral Si valence (e)
  P[OSHIFT + 1] = 6.0f;  // neutral O valence (e)
  if (hipMalloc(&m->d_P, sizeof(float) * NPARAM) != hipSuccess ||
      hipMalloc(&m->d_grad, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMalloc(&m->d_m, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMalloc(&m->d_v, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMalloc(&m->d_partial, sizeof(float) * NTRAIN * NPB) != hipSuccess ||
      hipMalloc(&m->d_loss, sizeof(double)) != hipSuccess) {
    free_params(m);
    delete m;
    return set_err(c, LTB_ERR_NOMEM, "MLP allocation failed");
  }
  if (hipMemcpy(m->d_P, P.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMemset(m->d_v, 0, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMemset(m->d_partial, 0, sizeof(float) * NTRAIN * NPB) != hipSuccess) {
    free_params(m);
    delete m;
    return set_err(c, LTB_ERR_DEVICE, "MLP parameter initialisation failed");
  }
  *out = m;
  return LTB_OK;
}

void ltb_mlp_destroy(ltb_mlp* m) {
  if (!m) return;
  (void)hipStreamSynchronize(m->ctx->stream);
  free_data(m);
  free_params(m);
  delete m;
}

int ltb_mlp_get_params(ltb_mlp* m, float* p) {
  if (!m || !p) return LTB_ERR_ARG;
  LTB_HIP(m->ctx, hipMemcpy(p, m->d_P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return LTB_OK;
}

int ltb_mlp_set_params(ltb_mlp* m, const float* p) {
  if (!m || !p) return LTB_ERR_ARG;
  LTB_HIP(m->ctx, hipMemcpy(m->d_P, p, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  m->adam_t = 0;
  LTB_HIP(m->ctx, hipMemset(m->d_m, 0, sizeof(float) * NTRAIN));
  LTB_HIP(m->ctx, hipMemset(m->d_v, 0, sizeof(float) * NTRAIN));
  return LTB_OK;
}

int ltb_descriptors(ltb_ctx* c, const ltb_system* s, float* feat) {
  if (!c || !s || !feat) return LTB_ERR_ARG;
  float* d = nullptr;
  LTB_HIP(c, hipMalloc(&d, sizeof(float) * NF * (size_t)s->natoms));
  int st = descriptors_device(c, s, d);
  if (!st) {
    hipError_t e = hipMemcpy(feat, d, sizeof(float) * NF * (size_t)s->natoms, hipMemcpyDeviceToHost);
    if (e != hipSuccess) st = set_err(c, LTB_ERR_DEVICE, hipGetErrorString(e));
  }
  (void)hipFree(d);
  return st;
}

int ltb_mlp_set_data(ltb_mlp* m, int n, const float* feat, const int* Z, int nseg, const int* seg,
                     const double* qtot, const double* target) {
  if (!m || n <= 0 || !feat || !Z || nseg <= 0 || !seg || !qtot) return LTB_ERR_ARG;
  ltb_ctx* c = m->ctx;
  if (seg[0] != 0 || seg[nseg] != n) return set_err(c, LTB_ERR_ARG, "segments must cover [0, n)");
  for (int s = 0; s < nseg; s++)
    if (seg[s + 1] <= seg[s]) return set_err(c, LTB_ERR_ARG, "empty segment");
  free_data(m);
  std::vector<int> idx[2];
  for (int a = 0; a < n; a++) {
    if (Z[a] != 14 && Z[a] != 8) return set_err(c, LTB_ERR_ARG, "Z must be 14 or 8");
    idx[Z[a] == 14 ? 0 : 1].push_back(a);
  }
  // n and nseg are set only once every buffer exists; any failure frees what
  // was allocated and leaves the object without data (n = 0)
  const int st = upload_data(m, n, feat, Z, nseg, seg, qtot, target, idx);
  if (st) {
    free_data(m);
    return st;
  }
  m->n = n;
  m->nseg = nseg;
  return LTB_OK;
}

}  // extern "C"

namespace {
int upload_data(ltb_mlp* m, int n, const float* feat, const int* Z, int nseg, const int* seg, const double* qtot,
                const double* target, const std::vector<int> (&idx)[2]) {
  ltb_ctx* c = m->ctx;
  LTB_HIP(c, hipMalloc(&m->d_feat, sizeof(float) * NF * (size_t)n));
  LTB_HIP(c, hipMalloc(&m->d_Z, sizeof(int) * n));
  LTB_HIP(c, hipMalloc(&m->d_seg, sizeof(int) * (nseg + 1)));
  for (int sp = 0; sp < 2; sp++) {
    m->nidx[sp] = (int)idx[sp].size();
    LTB_HIP(c, hipMalloc(&m->d_idx[sp], sizeof(int) * std::max(1, m->nidx[sp])));
    if (m->nidx[sp])
      LTB_HIP(c, hipMemcpy(m->d_idx[sp], idx[sp].data(), sizeof(int) * m->nidx[sp], hipMemcpyHostToDevice));
  }
  for (double** p : {&m->d_t, &m->d_qraw, &m->d_q, &m->d_g}) LTB_HIP(c, hipMalloc(p, sizeof(double) * n));
  LTB_HIP(c, hipMalloc(&m->d_qtot, sizeof(double) * nseg));
  LTB_HIP(c, hipMalloc(&m->d_lpart, sizeof(double) * nseg));
  LTB_HIP(c, hipMemcpy(m->d_feat, feat, sizeof(float) * NF * (size_t)n, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(m->d_Z, Z, sizeof(int) * n, hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(m->d_seg, seg, sizeof(int) * (nseg + 1), hipMemcpyHostToDevice));
  LTB_HIP(c, hipMemcpy(m->d_qtot, qtot, sizeof(double) * nseg, hipMemcpyHostToDevice));
  m->has_target = target != nullptr;
  if (target) LTB_HIP(c, hipMemcpy(m->d_t, target, sizeof(double) * n, hipMemcpyHostToDevice));
  return LTB_OK;
}
}  // namespace

extern "C" {

int ltb_mlp_fit_normalization(ltb_mlp* m) {
  if (!m || m->n <= 0 || !m->has_target) return LTB_ERR_ARG;
  ltb_ctx* c = m->ctx;
  std::vector<float> feat((size_t)NF * m->n), P(NPARAM);
  std::vector<double> t(m->n);
  std::vector<int> Z(m->n);
  LTB_HIP(c, hipMemcpy(feat.data(), m->d_feat, sizeof(float) * feat.size(), hipMemcpyDeviceToHost));
  LTB_HIP(c, hipMemcpy(t.data(), m->d_t, sizeof(double) * m->n, hipMemcpyDeviceToHost));
  LTB_HIP(c, hipMemcpy(Z.data(), m->d_Z, sizeof(int) * m->n, hipMemcpyDeviceToHost));
  LTB_HIP(c, hipMemcpy(P.data(), m->d_P, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  for (int sp = 0; sp < 2; sp++) {
    std::vector<double> s1(NF, 0.0), s2(NF, 0.0);
    double ts = 0;
    int cnt = 0;
    for (int a = 0; a < m->n; a++) {
      if ((Z[a] == 14 ? 0 : 1) != sp) continue;
      cnt++;
      ts += t[a];
      for (int f = 0; f < NF; f++) {
        double x = feat[(size_t)a * NF + f];
        s1[f] += x;
        s2[f] += x * x;
      }
    }
    if (!cnt) continue;
    for (int f = 0; f < NF; f++) {
      double mu = s1[f] / cnt, var = std::max(0.0, s2[f] / cnt - mu * mu);
      P[OMEAN + sp * NF + f] = (float)mu;
      P[OSD + sp * NF + f] = (float)std::max(std::sqrt(var), 1e-3);
    }
    P[OSHIFT + sp] = (float)(ts / cnt);
  }
  LTB_HIP(c, hipMemcpy(m->d_P, P.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  return LTB_OK;
}

int ltb_mlp_forward(ltb_mlp* m, double* q) {
  if (!m || m->n <= 0 || !q) return LTB_ERR_ARG;
  int st = forward_async(m);
  if (st) return st;
  LTB_HIP(m->ctx, hipMemcpyAsync(q, m->d_q, sizeof(double) * m->n, hipMemcpyDeviceToHost, m->ctx->stream));
  LTB_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  return LTB_OK;
}

int ltb_mlp_loss_grad(ltb_mlp* m, double* loss, float* grad) {
  if (!m || m->n <= 0 || !m->has_target) return LTB_ERR_ARG;
  int st;
  if ((st = forward_async(m)) || (st = loss_grad_async(m))) return st;
  if (grad)
    LTB_HIP(m->ctx, hipMemcpyAsync(grad, m->d_grad, sizeof(float) * NTRAIN, hipMemcpyDeviceToHost, m->ctx->stream));
  double l = 0;
  LTB_HIP(m->ctx, hipMemcpyAsync(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
  LTB_HIP(m->ctx, hipStreamSynchronize(m->ctx->stream));
  if (loss) *loss = l;
  return LTB_OK;
}

int ltb_mlp_adam(ltb_mlp* m, int steps, float lr, double* loss_hist) {
  if (!m || m->n <= 0 || !m->has_target || steps < 1 || !(lr > 0)) return LTB_ERR_ARG;
  ltb_ctx* c = m->ctx;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps; s++) {
    int st;
    if ((st = forward_async(m)) || (st = loss_grad_async(m))) return st;
    m->adam_t++;
    const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((NTRAIN + 255) / 256), dim3(256), 0, c->stream, m->d_P, m->d_grad,
                       m->d_m, m->d_v, lr, b1, b2, eps, c1, c2);
    LTB_HIP(c, hipGetLastError());
    if (loss_hist)
      LTB_HIP(c, hipMemcpyAsync(loss_hist + s, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  LTB_HIP(c, hipStreamSynchronize(c->stream));
  return LTB_OK;
}

int ltb_mlp_predict(ltb_mlp* m, const ltb_system* s, double* p_guess) {
  if (!m || !s || !p_guess) return LTB_ERR_ARG;
  ltb_ctx* c = m->ctx;
  std::vector<float> feat((size_t)NF * s->natoms);
  int st = ltb_descriptors(c, s, feat.data());
  if (st) return st;
  // a one-structure data set without targets (training data is dropped)
  std::vector<int> seg = {0, s->natoms};
  double q = s->nelectrons;
  if ((st = ltb_mlp_set_data(m, s->natoms, feat.data(), s->Z.data(), 1, seg.data(), &q, nullptr))) return st;
  std::vector<double> qa(s->natoms);
  if ((st = ltb_mlp_forward(m, qa.data()))) return st;
  for (int a = 0; a < s->natoms; a++) p_guess[a] = 0.5 * qa[a];
  return LTB_OK;
}

}  // extern "C"
