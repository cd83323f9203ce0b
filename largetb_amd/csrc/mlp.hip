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
#include <hip/hip_runtime.h>

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
constexpr int NPART = 64;  // atom partitions of the weight-gradient reduction

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
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

__global__ __launch_bounds__(128) void k_mlp_forward(int n, const float* __restrict__ feat, const int* __restrict__ Z,
                              const float* __restrict__ P, float* __restrict__ z1,
                              float* __restrict__ z2, float* __restrict__ z3,
                              double* __restrict__ qraw) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int c = species(Z[a]);
  const float* W = P + c * PS;
  float x[NF];
#pragma unroll
  for (int f = 0; f < NF; f++)
    x[f] = (feat[(size_t)a * NF + f] - P[OMEAN + c * NF + f]) / P[OSD + c * NF + f];
  float h[NH];
  for (int o = 0; o < NH; o++) {
    float s = W[OB1 + o];
#pragma unroll
    for (int f = 0; f < NF; f++) s = fmaf(W[OW1 + o * NF + f], x[f], s);
    z1[(size_t)a * NH + o] = s;
  }
#pragma unroll
  for (int f = 0; f < NH; f++) h[f] = silu(z1[(size_t)a * NH + f]);
  for (int o = 0; o < NH; o++) {
    float s = W[OB2 + o];
#pragma unroll
    for (int f = 0; f < NH; f++) s = fmaf(W[OW2 + o * NH + f], h[f], s);
    z2[(size_t)a * NH + o] = s;
  }
#pragma unroll
  for (int f = 0; f < NH; f++) h[f] = silu(z2[(size_t)a * NH + f]);
  for (int o = 0; o < NH; o++) {
    float s = W[OB3 + o];
#pragma unroll
    for (int f = 0; f < NH; f++) s = fmaf(W[OW3 + o * NH + f], h[f], s);
    z3[(size_t)a * NH + o] = s;
  }
  float y = W[OB4];
#pragma unroll
  for (int f = 0; f < NH; f++) y = fmaf(W[OW4 + f], silu(z3[(size_t)a * NH + f]), y);
  qraw[a] = (double)P[OSHIFT + c] + (double)y;
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

__global__ void k_sum(const double* __restrict__ v, int n, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double s = 0;
  for (int i = 0; i < n; i++) s += v[i];
  *out = s;
}

__global__ __launch_bounds__(128) void k_mlp_backward(int n, const int* __restrict__ Z, const float* __restrict__ P,
                               const float* __restrict__ z1, const float* __restrict__ z2,
                               const float* __restrict__ z3, const double* __restrict__ g,
                               float* __restrict__ d1, float* __restrict__ d2,
                               float* __restrict__ d3, float* __restrict__ d4) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const float* W = P + species(Z[a]) * PS;
  const float dy = (float)g[a];
  d4[a] = dy;
  float dz[NH];
#pragma unroll
  for (int o = 0; o < NH; o++) dz[o] = dy * W[OW4 + o] * dsilu(z3[(size_t)a * NH + o]);
#pragma unroll
  for (int o = 0; o < NH; o++) d3[(size_t)a * NH + o] = dz[o];
  // d2 = (W3^T dz3) * silu'(z2)
  for (int f = 0; f < NH; f++) {
    float s = 0;
#pragma unroll
    for (int o = 0; o < NH; o++) s = fmaf(W[OW3 + o * NH + f], dz[o], s);
    d2[(size_t)a * NH + f] = s * dsilu(z2[(size_t)a * NH + f]);
  }
#pragma unroll
  for (int o = 0; o < NH; o++) dz[o] = d2[(size_t)a * NH + o];
  for (int f = 0; f < NH; f++) {
    float s = 0;
#pragma unroll
    for (int o = 0; o < NH; o++) s = fmaf(W[OW2 + o * NH + f], dz[o], s);
    d1[(size_t)a * NH + f] = s * dsilu(z1[(size_t)a * NH + f]);
  }
}

// partial[p][...] = sum over atoms of species c in partition p of D[a][o] * Hin[a][i]
// layer 1: Hin = normalized features (NF); 2: silu(z1); 3: silu(z2); 4: silu(z3), D = d4
__global__ __launch_bounds__(128) void k_wgrad(int layer, int c, const int* __restrict__ idx, int nidx,
                                               const float* __restrict__ feat,
                                               const float* __restrict__ P,
                                               const float* __restrict__ hz,  // z of the input layer
                                               const float* __restrict__ D, float* __restrict__ partial) {
  const int o = blockIdx.x, p = blockIdx.y, i = threadIdx.x;
  const int nin = layer == 1 ? NF : NH;
  const int nout = layer == 4 ? 1 : NH;
  if (i > nin) return;
  const int per = (nidx + NPART - 1) / NPART;
  const int ab = p * per, ae = min(nidx, ab + per);
  float acc = 0;
  for (int t = ab; t < ae; t++) {
    const int a = idx[t];
    const float d = D[(size_t)a * nout + o];
    float h;
    if (i == nin)
      h = 1.0f;  // bias
    else if (layer == 1)
      h = (feat[(size_t)a * NF + i] - P[OMEAN + c * NF + i]) / P[OSD + c * NF + i];
    else
      h = silu(hz[(size_t)a * NH + i]);
    acc = fmaf(d, h, acc);
  }
  int woff, boff;
  if (layer == 1) woff = OW1, boff = OB1;
  else if (layer == 2) woff = OW2, boff = OB2;
  else if (layer == 3) woff = OW3, boff = OB3;
  else woff = OW4, boff = OB4;
  float* out = partial + (size_t)p * NTRAIN + c * PS;
  if (i == nin)
    out[boff + o] = acc;
  else
    out[woff + o * nin + i] = acc;
}

__global__ void k_reduce_partials(const float* __restrict__ partial, float* __restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= NTRAIN) return;
  float s = 0;
  for (int p = 0; p < NPART; p++) s += partial[(size_t)p * NTRAIN + j];
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
  float *d_feat = nullptr, *d_z1 = nullptr, *d_z2 = nullptr, *d_z3 = nullptr;
  float *d_d1 = nullptr, *d_d2 = nullptr, *d_d3 = nullptr, *d_d4 = nullptr;
  int *d_Z = nullptr, *d_seg = nullptr, *d_idx[2] = {nullptr, nullptr};
  int nidx[2] = {0, 0};
  double *d_qtot = nullptr, *d_t = nullptr, *d_qraw = nullptr, *d_q = nullptr, *d_g = nullptr,
         *d_lpart = nullptr, *d_loss = nullptr;
  bool has_target = false;
};

namespace {

void free_data(ltb_mlp* m) {
  void* ps[] = {m->d_feat, m->d_z1, m->d_z2, m->d_z3, m->d_d1, m->d_d2, m->d_d3, m->d_d4, m->d_Z,
                m->d_seg, m->d_idx[0], m->d_idx[1], m->d_qtot, m->d_t, m->d_qraw, m->d_q, m->d_g,
                m->d_lpart};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->d_feat = m->d_z1 = m->d_z2 = m->d_z3 = m->d_d1 = m->d_d2 = m->d_d3 = m->d_d4 = nullptr;
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

int forward_async(ltb_mlp* m) {
  ltb_ctx* c = m->ctx;
  const int g = (m->n + 127) / 128;
  hipLaunchKernelGGL(k_mlp_forward, dim3(g), dim3(128), 0, c->stream, m->n, m->d_feat, m->d_Z, m->d_P,
                     m->d_z1, m->d_z2, m->d_z3, m->d_qraw);
  hipLaunchKernelGGL(k_project, dim3(m->nseg), dim3(256), 0, c->stream, m->d_seg, m->d_qtot, m->d_qraw,
                     m->d_q);
  LTB_HIP(c, hipGetLastError());
  return LTB_OK;
}

int loss_grad_async(ltb_mlp* m) {
  ltb_ctx* c = m->ctx;
  hipLaunchKernelGGL(k_loss_grad, dim3(m->nseg), dim3(256), 0, c->stream, m->d_seg, m->d_q, m->d_t,
                     m->n, m->d_g, m->d_lpart);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(64), 0, c->stream, m->d_lpart, m->nseg, m->d_loss);
  const int g = (m->n + 127) / 128;
  hipLaunchKernelGGL(k_mlp_backward, dim3(g), dim3(128), 0, c->stream, m->n, m->d_Z, m->d_P, m->d_z1,
                     m->d_z2, m->d_z3, m->d_g, m->d_d1, m->d_d2, m->d_d3, m->d_d4);
  for (int sp = 0; sp < 2; sp++) {
    if (m->nidx[sp] == 0) continue;
    hipLaunchKernelGGL(k_wgrad, dim3(NH, NPART), dim3(128), 0, c->stream, 1, sp, m->d_idx[sp],
                       m->nidx[sp], m->d_feat, m->d_P, (const float*)nullptr, m->d_d1, m->d_partial);
    hipLaunchKernelGGL(k_wgrad, dim3(NH, NPART), dim3(128), 0, c->stream, 2, sp, m->d_idx[sp],
                       m->nidx[sp], m->d_feat, m->d_P, m->d_z1, m->d_d2, m->d_partial);
    hipLaunchKernelGGL(k_wgrad, dim3(NH, NPART), dim3(128), 0, c->stream, 3, sp, m->d_idx[sp],
                       m->nidx[sp], m->d_feat, m->d_P, m->d_z2, m->d_d3, m->d_partial);
    hipLaunchKernelGGL(k_wgrad, dim3(1, NPART), dim3(128), 0, c->stream, 4, sp, m->d_idx[sp],
                       m->nidx[sp], m->d_feat, m->d_P, m->d_z3, m->d_d4, m->d_partial);
  }
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
      hipMalloc(&m->d_partial, sizeof(float) * NTRAIN * NPART) != hipSuccess ||
      hipMalloc(&m->d_loss, sizeof(double)) != hipSuccess) {
    free_params(m);
    delete m;
    return set_err(c, LTB_ERR_NOMEM, "MLP allocation failed");
  }
  if (hipMemcpy(m->d_P, P.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMemset(m->d_v, 0, sizeof(float) * NTRAIN) != hipSuccess ||
      hipMemset(m->d_partial, 0, sizeof(float) * NTRAIN * NPART) != hipSuccess) {
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
  const size_t act = sizeof(float) * NH * (size_t)n;
  LTB_HIP(c, hipMalloc(&m->d_feat, sizeof(float) * NF * (size_t)n));
  for (float** p : {&m->d_z1, &m->d_z2, &m->d_z3, &m->d_d1, &m->d_d2, &m->d_d3}) LTB_HIP(c, hipMalloc(p, act));
  LTB_HIP(c, hipMalloc(&m->d_d4, sizeof(float) * n));
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
