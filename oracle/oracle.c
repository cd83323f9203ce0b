/*
 * oracle.c -- CPU FP64 restatement of pghale/largeTB's KPE / SP2 / SCC path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  It is the checker the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.  The product (largetb_amd/) never links it.
 *
 * The reference itself (withSP2_deltaH.cpp) cannot be compiled here: it needs
 * PETSc (petscksp.h), which is not installed and cannot be installed offline.
 * This file restates its algorithm in plain C, function by function.
 *
 * Pins (checks against something other than this file's own formulas; the
 * tests live in tests/test_oracle_*.py):
 *   orc_quartz + orc_neighbors + orc_sk_block + orc_build_h
 *        -> Amat.m (1 cell) and H.m (4x4x4) in reference mode, every entry.
 *   orc_build_gamma -> GAMMA_AB.m; Gamma . netCharge.m -> GammaQ_vector.m.
 *   orc_gershgorin -> brute force over the dense H.m rows, and the spectrum
 *        (numpy eigvalsh) must lie inside the bounds.
 *   orc_jackson -> gnsfile.txt column 1;  orc_cheb_nodes -> dosre.txt column 1.
 *   orc_gmM -> closed form: g_m is the autocorrelation of a_i; its cosine
 *        series is non-negative (Fejer-Riesz), g_{M-1} = 0.
 *   orc_step_coeffs -> kT -> 0 limit of orc_fd_coeffs, and the Chebyshev
 *        expansion of the step reproduces theta(mu - x) away from mu.
 *   orc_fd_coeffs -> quadrature convergence (K vs 2K), and the truncated
 *        series evaluated at points reproduces the Fermi function.
 *   orc_kpe_diag_moments -> [T_n]_jj = sum_k v_jk^2 T_n(e_k) from numpy eigh
 *        (brute force on 72 atoms), and sum over j = trace identity.
 *   orc_kpe_apply -> dense polynomial sum_m w_m T_m(H~) x built by numpy.
 *   orc_density (mu, populations, band energy) -> the same quantities from
 *        eigenpairs with the KPE-damped occupations; sum pop = Ne/2.
 *   orc_scc -> fixed point q = F(q) checked with an independent numpy
 *        eigen-solver SCC; result independent of the starting guess.
 *   orc_sp2_dense -> eigen-projector onto the Ne/2 lowest states.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double PI_REF = 3.141592653589793238463; /* deltaH:28 */
static const double ANG_TO_BOHR = 1.88973;            /* deltaH:507 */

/* ------------------------------------------------------------------ */
/* random numbers: SplitMix64 -> Box-Muller (project-defined, portable) */
uint64_t orc_splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double u01(uint64_t *s) { return (double)((orc_splitmix64(s) >> 11) + 1) * 0x1.0p-53; }

double orc_normal(uint64_t *s, int *have, double *spare) {
  if (*have) {
    *have = 0;
    return *spare;
  }
  double u1 = u01(s), u2 = u01(s);
  double r = sqrt(-2.0 * log(u1));
  double t = 2.0 * PI_REF * u2;
  *spare = r * sin(t);
  *have = 1;
  return r * cos(t);
}

/* ------------------------------------------------------------------ */
/* geometry: initializeCoordinates deltaH:504-667 (fractional coordinates
 * of the 9 atoms: :530-654; note the second Si's z = 0.6777, :546-548)   */
static const double FRAC[9][3] = {
    {0.4697, 0.0000, 0.0000}, {0.0000, 0.4697, 0.6777}, {0.5303, 0.5303, 0.3333},
    {0.4133, 0.2672, 0.1188}, {0.2672, 0.4133, 0.5479}, {0.7328, 0.1461, 0.7855},
    {0.5867, 0.8539, 0.2145}, {0.8539, 0.5867, 0.4521}, {0.1461, 0.7328, 0.8812}};
static const int ZCELL[9] = {14, 14, 14, 8, 8, 8, 8, 8, 8};

int orc_quartz(int L, int M, int N, int mode, double sigma_ang, uint64_t seed, double *xyz,
               int *Z, double *cell) {
  if (L <= 0 || M <= 0 || N <= 0) return ORC_ERR_ARG;
  double a, c;
  if (mode == ORC_MODE_REFERENCE) {
    a = 9.725; /* deltaH:508 */
    c = 10.7;
  } else {
    a = 4.913 * ANG_TO_BOHR;
    c = 5.405 * ANG_TO_BOHR;
  }
  double R[3][3] = {{a, 0.0, 0.0}, {-a * sin(PI_REF / 6), a * cos(PI_REF / 6), 0.0}, {0.0, 0.0, c}};
  int at = 0;
  for (int i = 0; i < L; i++)
    for (int j = 0; j < M; j++)
      for (int k = 0; k < N; k++) {
        double o[3];
        for (int d = 0; d < 3; d++) o[d] = i * R[0][d] + j * R[1][d] + k * R[2][d];
        for (int s = 0; s < 9; s++, at++) {
          for (int d = 0; d < 3; d++)
            xyz[3 * at + d] = o[d] + (FRAC[s][0] * R[0][d] + FRAC[s][1] * R[1][d] + FRAC[s][2] * R[2][d]);
          Z[at] = ZCELL[s];
        }
      }
  if (mode == ORC_MODE_CONFIG && sigma_ang > 0) {
    uint64_t st = seed;
    int have = 0;
    double spare = 0;
    for (int i = 0; i < 3 * at; i++) xyz[i] += sigma_ang * ANG_TO_BOHR * orc_normal(&st, &have, &spare);
  }
  if (cell) {
    for (int d = 0; d < 3; d++) {
      cell[0 * 3 + d] = L * R[0][d];
      cell[1 * 3 + d] = M * R[1][d];
      cell[2 * 3 + d] = N * R[2][d];
    }
  }
  return ORC_OK;
}

/* getHaloAtoms deltaH:372-501: every image (lp,mp,np) in order, every atom j;
 * neighbour iff 0.1 < r^2 < 16 Bohr^2 (:392, :421-422); at most 20 (:357). */
int orc_neighbors(int nat, const double *xyz, const double *cell, int pmask, int *nn, int *nbr,
                  double *nd) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    lo[d] = (pmask >> d & 1) ? -1 : 0;
    hi[d] = (pmask >> d & 1) ? 1 : 0;
  }
  int status = ORC_OK;
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < nat; i++) {
    int cnt = 0;
    for (int l = lo[0]; l <= hi[0]; l++)
      for (int m = lo[1]; m <= hi[1]; m++)
        for (int n = lo[2]; n <= hi[2]; n++) {
          double sh[3];
          for (int d = 0; d < 3; d++) sh[d] = l * cell[d] + m * cell[3 + d] + n * cell[6 + d];
          for (int j = 0; j < nat; j++) {
            double dx = xyz[3 * j] + sh[0] - xyz[3 * i];
            double dy = xyz[3 * j + 1] + sh[1] - xyz[3 * i + 1];
            double dz = xyz[3 * j + 2] + sh[2] - xyz[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < 16.0 && r2 > 0.1) {
              if (cnt >= ORC_MAXN) {
#pragma omp atomic write
                status = ORC_ERR_NEIGHBORS;
                continue;
              }
              nbr[i * ORC_MAXN + cnt] = j;
              nd[(i * ORC_MAXN + cnt) * 3 + 0] = dx;
              nd[(i * ORC_MAXN + cnt) * 3 + 1] = dy;
              nd[(i * ORC_MAXN + cnt) * 3 + 2] = dz;
              cnt++;
            }
          }
        }
    nn[i] = cnt;
  }
  return status;
}

/* ------------------------------------------------------------------ */
/* Slater-Koster tables, deltaH:75-142 (columns cp2, cp3, cp5, cp6)      */
static const double SISI[20][4] = {
    {17.865, -5.7398, 17.262, -13.794},   {87.114, 5.8664, 39.685, -7.1328},
    {-722.33, -52.111, -283.25, 27.816},  {66.059, -3.2651, 34.498, -27.631},
    {-613.94, -6.6631, -41.298, 12.075},  {-727.27, 1.3757, -112.62, 5.9498},
    {246.1, -84.273, 241.71, -164.58},    {-1564.3, 4.1541, -179.83, 22.449},
    {-589.53, -7.9678, 215.3, -36.682},   {-1020.8, 33.73, 7.8809, -87.911},
    {-812.85, -33.408, 13.727, 158.22},   {1189.7, 49.751, 54.842, -56.155},
    {-96.931, 8.9531, 102.98, -91.347},   {860.35, 1.2204, -44.669, 12.487},
    {3551.6, 9.5425, -32.771, 10.438},    {468.97, 1.6056, -128.73, -1.5888},
    {-317.44, 6.9669, 133.21, 82.532},    {517.74, 47.052, -56.922, 1.1376},
    {210.8, -18.241, -18.441, 80.163},    {-443.57, 36.135, 52.835, -13.413}};
static const double OO[20][4] = {
    {5.5526, -0.82118, 3.813, -1.2323},   {-31.886, 10.518, -41.071, 52.454},
    {233.96, -39.064, 272.9, -407.19},    {-60.015, 61.253, -128.95, -294.22},
    {213.73, -204.34, 410.72, 1263.6},    {49.054, 32.239, -78.906, -1060},
    {-73.081, 132.93, -81.269, 1956.2},   {33.537, -13.278, -137.34, -384.47},
    {78.858, 165.14, -273.61, -873.53},   {-167.41, 68.088, -277.49, 642.65},
    {75.731, 7.9468, -767.04, -615.24},   {-63.795, -44.453, -288.57, 988.73},
    {-78.765, 32.997, -76.092, -811.61},  {124.63, -144.7, -189.3, 554.51},
    {40.615, -149.44, 780.52, 379.23},    {82.116, -29.881, 410.4, -93.407},
    {-80.656, -54.736, 203.1, -131.08},   {-74.163, 16.814, 710.57, -213.88},
    {7.3441, 20.02, -100.83, 128.52},     {-93.219, 6.9584, 267.84, -242.74}};
static const double SIO[20][4] = {
    {13.778, -2.7061, 8.6477, -7.6931},   {9.0798, 19.494, -22.43, 52.885},
    {-15.173, -145.38, 262.06, -468.23},  {9.4007, -23.026, 223.86, -100.16},
    {40.783, 126.17, -785.64, 228.63},    {23.249, -34.501, 334.88, -253.56},
    {-17.829, -75.476, 156.18, -638.68},  {18.37, -22.003, 10.705, -312.46},
    {1.7289, 75.972, 572.55, -1010.6},    {-30.424, 17.518, -55.884, -478.46},
    {-31.663, -36.928, -349.49, -688.24}, {-82.425, -26.609, -390.25, -296.52},
    {35.07, 77.948, -143.41, 157.81},     {-27.066, -22.899, -390.79, 580.81},
    {-20.555, -140.91, -152.48, 401.59},  {49.585, -5.9967, 502.82, 768.76},
    {-8.7806, 109.97, -113.52, 980.42},   {37.13, -18.019, -148.42, 243.34},
    {-20.674, 15.652, 904.76, 317.15},    {16.513, 39.501, -630.1, 84.78}};

/* get4by4HmatrixElements deltaH:211-349 */
int orc_sk_block(int zi, int zj, double x, double y, double z, double *h) {
  double r2 = x * x + y * y + z * z;
  double r = sqrt(r2);
  const double(*tab)[4] = NULL;
  for (int k = 0; k < 16; k++) h[k] = 0.0;
  if ((zi == 14 && zj == 8) || (zi == 8 && zj == 14)) {
    if (r < 0.01) return ORC_ERR_GEOMETRY; /* :219-222 */
    tab = SIO;
  } else if (zi == 14 && zj == 14) {
    if (r < 0.01) { /* onsite, :258-263 */
      h[0] = -0.39572506;
      h[5] = h[10] = h[15] = -0.15031380;
      return ORC_OK;
    }
    tab = SISI;
  } else if (zi == 8 && zj == 8) {
    if (r < 0.01) { /* onsite, :299-304 */
      h[0] = -0.87883246;
      h[5] = h[10] = h[15] = -0.33213167;
      return ORC_OK;
    }
    tab = OO;
  } else {
    return ORC_ERR_GEOMETRY; /* unknown pair, :338-341 */
  }
  /* radial sums S_c = sum_p tab[p][c] * T_p(r/10) * exp(-p r) */
  double S[4] = {0, 0, 0, 0};
  for (int p = 1; p <= 20; p++) {
    double f = cos(p * acos(r / 10)) * exp(-p * r);
    for (int c = 0; c < 4; c++) S[c] += tab[p - 1][c] * f;
  }
  const double d[3] = {x, y, z};
  h[0] = S[3]; /* ss: cp6 */
  for (int a = 0; a < 3; a++) {
    h[0 * 4 + 1 + a] = S[2] * d[a] / r;  /* sp: +cp5 d/r  (:225-227) */
    h[(1 + a) * 4 + 0] = -S[2] * d[a] / r; /* ps: -cp5 d/r  (:230) */
    for (int b = 0; b < 3; b++) {
      double dd = d[a] * d[b] / r2;
      h[(1 + a) * 4 + 1 + b] = S[0] * dd - S[1] * (dd - (a == b ? 1.0 : 0.0));
    }
  }
  return ORC_OK;
}

/* constructH deltaH:969-1066 as BSR with INSERT_VALUES semantics */
int orc_build_h(int nat, const int *Z, const int *nn, const int *nbr, const double *nd, int cap,
                int *rowptr, int *col, double *val) {
  int nz = 0;
  rowptr[0] = 0;
  for (int i = 0; i < nat; i++) {
    int start = nz;
    if (nz >= cap) return -ORC_ERR_ARG;
    col[nz] = i;
    int st = orc_sk_block(Z[i], Z[i], 0.0, 0.0, 0.0, val + 16 * (size_t)nz);
    if (st) return -st;
    nz++;
    for (int k = 0; k < nn[i]; k++) {
      int j = nbr[i * ORC_MAXN + k];
      const double *d = nd + (size_t)(i * ORC_MAXN + k) * 3;
      int slot = -1;
      for (int s = start; s < nz; s++)
        if (col[s] == j) slot = s;
      if (slot < 0) {
        if (nz >= cap) return -ORC_ERR_ARG;
        slot = nz++;
        col[slot] = j;
      }
      st = orc_sk_block(Z[i], Z[j], d[0], d[1], d[2], val + 16 * (size_t)slot);
      if (st) return -st;
    }
    rowptr[i + 1] = nz;
  }
  return nz;
}

/* initializeForSCF deltaH:854-900 */
int orc_build_gamma(int nat, const int *nn, const int *nbr, const double *nd, double Us, int cap,
                    int *rowptr, int *col, double *val) {
  int nz = 0;
  rowptr[0] = 0;
  for (int i = 0; i < nat; i++) {
    int start = nz;
    if (nz >= cap) return -ORC_ERR_ARG;
    col[nz] = i;
    val[nz] = -Us;
    nz++;
    for (int k = 0; k < nn[i]; k++) {
      int j = nbr[i * ORC_MAXN + k];
      const double *d = nd + (size_t)(i * ORC_MAXN + k) * 3;
      double r = pow(d[0] * d[0] + d[1] * d[1] + d[2] * d[2], 0.5);
      int slot = -1;
      for (int s = start; s < nz; s++)
        if (col[s] == j) slot = s;
      if (slot < 0) {
        if (nz >= cap) return -ORC_ERR_ARG;
        slot = nz++;
        col[slot] = j;
      }
      val[slot] = 1 / r;
    }
    rowptr[i + 1] = nz;
  }
  return nz;
}

/* GershGorinParallel deltaH:144-209 */
int orc_gershgorin(int nat, const int *rowptr, const int *col, const double *val,
                   const double *shift, double *emin, double *emax) {
  double lo = INFINITY, hi = -INFINITY;
  int bad = 0;
  for (int i = 0; i < nat; i++)
    for (int a = 0; a < 4; a++) {
      double dg = 0, off = 0;
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
        for (int b = 0; b < 4; b++) {
          double v = val[16 * (size_t)k + a * 4 + b];
          if (!isfinite(v)) bad = 1;
          if (col[k] == i && b == a)
            dg = v;
          else
            off += fabs(v);
        }
      if (shift) dg += shift[4 * i + a];
      if (!isfinite(dg)) bad = 1;
      if (dg - off < lo) lo = dg - off;
      if (dg + off > hi) hi = dg + off;
    }
  *emin = lo;
  *emax = hi;
  return bad ? ORC_ERR_NONFINITE : ORC_OK;
}

/* ------------------------------------------------------------------ */
/* kernels and coefficients                                            */

/* getKPE gmMs, deltaH:1105-1132 */
void orc_gmM(int M, double *g) {
  double *U = (double *)malloc(sizeof(double) * M);
  double alpha = PI_REF / (M + 2);
  double lam = cos(alpha);
  double nrm = 0;
  for (int m = 0; m < M; m++) {
    U[m] = sin((m + 1) * acos(lam)) / sin(acos(lam));
    nrm += U[m] * U[m];
  }
  nrm = sqrt(nrm);
  for (int i = 0; i < M; i++) U[i] /= nrm;
  for (int m = 0; m < M; m++) {
    double s = 0;
    for (int i = 0; i < M - m - 1; i++) s += U[i] * U[m + i];
    g[m] = s;
  }
  free(U);
}

/* runKPM Jackson kernel, deltaH:1593-1595 */
void orc_jackson(int N, double *g) {
  for (int n = 0; n < N; n++) {
    double t = PI_REF * (double)n / (double)(N + 1);
    g[n] = ((N - n + 1) * cos(t) + sin(t) * (cos(PI_REF / (N + 1)) / sin(PI_REF / (N + 1)))) /
           (double)(N + 1);
  }
}

/* runKPM nodes, deltaH:1589-1591 */
void orc_cheb_nodes(int n, double *x) {
  for (int k = 0; k < n; k++) x[k] = cos(PI_REF * (k + 1.0 / 2.0) / n);
}

/* getKPE mu_mvals, deltaH:1100-1103 (step function at mu) */
void orc_step_coeffs(int M, double mu, double *c) {
  c[0] = 2 * (1 - acos(mu) / PI_REF);
  for (int m = 1; m < M; m++) c[m] = -2 * sin(m * acos(mu)) / (m * PI_REF);
}

static double fermi(double x, double mu, double kT) {
  double t = (x - mu) / kT;
  if (t > 0) {
    double e = exp(-t);
    return e / (1.0 + e);
  }
  return 1.0 / (1.0 + exp(t));
}

/* Chebyshev coefficients of the Fermi function by Chebyshev-Gauss
 * quadrature on K nodes: c_m = (2/K) sum_k f(cos th_k) cos(m th_k). */
void orc_fd_coeffs(int M, double mu, double kT, int K, double *c) {
  for (int m = 0; m < M; m++) c[m] = 0;
  for (int k = 0; k < K; k++) {
    double th = PI_REF * (k + 0.5) / K;
    double f = fermi(cos(th), mu, kT);
    for (int m = 0; m < M; m++) c[m] += f * cos(m * th);
  }
  for (int m = 0; m < M; m++) c[m] *= 2.0 / K;
}

/* ------------------------------------------------------------------ */
/* runKPM rescale deltaH:1554-1560: H~ = (H - b)/a                     */
int orc_rescaled_csr(int nat, const int *rowptr, const int *col, const double *val,
                     const double *shift, double a, double b, int *crow, int *ccol, double *cval) {
  int nz = 0;
  crow[0] = 0;
  for (int i = 0; i < nat; i++)
    for (int aa = 0; aa < 4; aa++) {
      for (int k = rowptr[i]; k < rowptr[i + 1]; k++)
        for (int bb = 0; bb < 4; bb++) {
          double v = val[16 * (size_t)k + aa * 4 + bb];
          if (col[k] == i && aa == bb) v = v + (shift ? shift[4 * i + aa] : 0.0) - b;
          ccol[nz] = 4 * col[k] + bb;
          cval[nz] = v / a;
          nz++;
        }
      crow[4 * i + aa + 1] = nz;
    }
  return nz;
}

static void spmv(int n, const int *rp, const int *ci, const double *v, const double *x, double *y) {
  for (int r = 0; r < n; r++) {
    double s = 0;
    for (int k = rp[r]; k < rp[r + 1]; k++) s += v[k] * x[ci[k]];
    y[r] = s;
  }
}
static void vcopy(int n, const double *x, double *y) { memcpy(y, x, sizeof(double) * n); }
static void vscale(int n, double *x, double s) {
  for (int i = 0; i < n; i++) x[i] *= s;
}
static void vaxpy(int n, double *y, double a, const double *x) {
  for (int i = 0; i < n; i++) y[i] += a * x[i];
}

/* getKPE's recursion (deltaH:1139-1168) on e_j, reading [T_n]_jj into
 * out[0..nm); xi, xm1, xp1, hx are n-vectors of scratch */
static void diag_moments_one(int n, const int *crow, const int *ccol, const double *cval, int nm, int j,
                             double *out, double *xi, double *xm1, double *xp1, double *hx) {
  memset(xi, 0, sizeof(double) * n);
  xi[j] = 1.0;                /* VecCopy(randomVector, x_i)         */
  out[0] = xi[j];
  spmv(n, crow, ccol, cval, xi, hx); /* MatMult(Htotal, x_i, Hx_i)  */
  vcopy(n, xi, xm1);                 /* VecCopy(x_i, x_m1)          */
  vcopy(n, hx, xi);                  /* VecCopy(Hx_i, x_i)          */
  out[1] = xi[j];
  for (int m = 2; m < nm; m++) {
    spmv(n, crow, ccol, cval, xi, hx); /* MatMult                  */
    vcopy(n, hx, xp1);                 /* VecCopy(Hx_i, x_p1)       */
    vscale(n, xp1, 2.0);               /* VecScale(x_p1, 2.0)       */
    vaxpy(n, xp1, -1.0, xm1);          /* VecAXPY(x_p1, -1, x_m1)   */
    vcopy(n, xi, xm1);                 /* VecCopy(x_i, x_m1)        */
    vcopy(n, xp1, xi);                 /* VecCopy(x_p1, x_i)        */
    out[m] = xi[j];                    /* theta-AXPY -> moment read */
  }
}

/* the columns cols[0..ncols) (or j0 + k when cols is NULL) */
static int diag_moments_list(int n, const int *crow, const int *ccol, const double *cval, int nm, int ncols,
                             int j0, const int *cols, double *mom, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
#pragma omp parallel num_threads(nthreads)
  {
    double *xi = (double *)malloc(sizeof(double) * n);
    double *xm1 = (double *)malloc(sizeof(double) * n);
    double *xp1 = (double *)malloc(sizeof(double) * n);
    double *hx = (double *)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 1)
    for (int k = 0; k < ncols; k++)
      diag_moments_one(n, crow, ccol, cval, nm, cols ? cols[k] : j0 + k, mom + (size_t)k * nm, xi, xm1, xp1,
                       hx);
    free(xi);
    free(xm1);
    free(xp1);
    free(hx);
  }
  return ORC_OK;
}

int orc_kpe_diag_moments(int n, const int *crow, const int *ccol, const double *cval, int nm,
                         int j0, int j1, double *mom, int nthreads) {
  if (nm < 2 || j0 < 0 || j1 > n || j0 > j1) return ORC_ERR_ARG;
  return diag_moments_list(n, crow, ccol, cval, nm, j1 - j0, j0, NULL, mom, nthreads);
}

int orc_kpe_diag_moments_cols(int n, const int *crow, const int *ccol, const double *cval, int nm,
                              int ncols, const int *cols, double *mom, int nthreads) {
  if (nm < 2 || ncols < 0 || !cols) return ORC_ERR_ARG;
  for (int k = 0; k < ncols; k++)
    if (cols[k] < 0 || cols[k] >= n) return ORC_ERR_ARG;
  return diag_moments_list(n, crow, ccol, cval, nm, ncols, 0, cols, mom, nthreads);
}

/* getKPE deltaH:1068-1185 with coefficients w */
int orc_kpe_apply(int n, const int *crow, const int *ccol, const double *cval, int M,
                  const double *w, const double *x, double *out) {
  double *xi = (double *)malloc(sizeof(double) * n);
  double *xm1 = (double *)malloc(sizeof(double) * n);
  double *xp1 = (double *)malloc(sizeof(double) * n);
  double *hx = (double *)malloc(sizeof(double) * n);
  double *th = (double *)calloc(n, sizeof(double));
  vcopy(n, x, xi);
  vaxpy(n, th, w[0], xi);
  spmv(n, crow, ccol, cval, xi, hx);
  vcopy(n, xi, xm1);
  vcopy(n, hx, xi);
  if (M > 1) vaxpy(n, th, w[1], xi);
  for (int i = 2; i < M; i++) {
    spmv(n, crow, ccol, cval, xi, hx);
    vcopy(n, hx, xp1);
    vscale(n, xp1, 2.0);
    vaxpy(n, xp1, -1.0, xm1);
    vcopy(n, xi, xm1);
    vcopy(n, xp1, xi);
    vaxpy(n, th, w[i], xi);
  }
  vcopy(n, th, out);
  free(xi);
  free(xm1);
  free(xp1);
  free(hx);
  free(th);
  return ORC_OK;
}

double orc_count(int M, const double *w, const double *mu) {
  double s = 0;
  for (int m = 0; m < M; m++) s += w[m] * mu[m];
  return s;
}

/* weights w_m = c_m g_m (m>=1), w_0 = c_0/2 -- getKPE :1136, :1142, :1166 */
static void kpe_weights(int M, const double *c, const double *g, double *w) {
  w[0] = c[0] / 2.0;
  for (int m = 1; m < M; m++) w[m] = c[m] * g[m];
}

static void coeffs_at(const orc_kpe_params *p, double mu, double kTs, double *c) {
  if (p->kT <= 0)
    orc_step_coeffs(p->M, mu, c);
  else
    orc_fd_coeffs(p->M, mu, kTs, p->K, c);
}

/* mu bisection on the exact moments (replaces runKPM's DOS search,
 * deltaH:1663-1685).  For Fermi-Dirac the count is evaluated as
 * sum_k f(x_k) W_k, W_k = (1/K)(mu_0 + 2 sum_{m>=1} g_m mu_m cos(m th_k)),
 * which is sum_m w_m mu_m with the quadrature coefficients (finite sums
 * exchanged). */
static double bisect_mu(const orc_kpe_params *p, double kTs, const double *g, const double *mu) {
  int M = p->M, K = p->K;
  double lo = -1.0, hi = 1.0;
  int iters = p->bisect_iters > 0 ? p->bisect_iters : 100;
  if (p->kT > 0) {
    double *W = (double *)malloc(sizeof(double) * K);
    double *X = (double *)malloc(sizeof(double) * K);
    for (int k = 0; k < K; k++) {
      double th = PI_REF * (k + 0.5) / K;
      double s = mu[0];
      for (int m = 1; m < M; m++) s += 2.0 * g[m] * mu[m] * cos(m * th);
      W[k] = s / K;
      X[k] = cos(th);
    }
    for (int it = 0; it < iters; it++) {
      double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      double n = 0;
      for (int k = 0; k < K; k++) n += fermi(X[k], mid, kTs) * W[k];
      if (2.0 * n < p->Ne)
        lo = mid;
      else
        hi = mid;
    }
    free(W);
    free(X);
  } else {
    double *c = (double *)malloc(sizeof(double) * M);
    double *w = (double *)malloc(sizeof(double) * M);
    for (int it = 0; it < iters; it++) {
      double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      orc_step_coeffs(M, mid, c);
      kpe_weights(M, c, g, w);
      if (2.0 * orc_count(M, w, mu) < p->Ne)
        lo = mid;
      else
        hi = mid;
    }
    free(c);
    free(w);
  }
  return 0.5 * (lo + hi);
}

int orc_density(int nat, const int *rowptr, const int *col, const double *val,
                const double *shift, const orc_kpe_params *p, double *pop, double *mom_atom,
                orc_kpe_result *res, int nthreads) {
  int n = 4 * nat, M = p->M, nm = M + 1;
  double emin, emax;
  int st = orc_gershgorin(nat, rowptr, col, val, shift, &emin, &emax);
  if (st) return st;
  double a = (emax - emin) / 2, b = (emax + emin) / 2;
  int nnz = rowptr[nat] * 16;
  int *crow = (int *)malloc(sizeof(int) * (n + 1));
  int *ccol = (int *)malloc(sizeof(int) * nnz);
  double *cval = (double *)malloc(sizeof(double) * nnz);
  orc_rescaled_csr(nat, rowptr, col, val, shift, a, b, crow, ccol, cval);
  double *mo = (double *)malloc(sizeof(double) * (size_t)n * nm);
  orc_kpe_diag_moments(n, crow, ccol, cval, nm, 0, n, mo, nthreads);
  double *ma = mom_atom ? mom_atom : (double *)malloc(sizeof(double) * (size_t)nat * nm);
  double *mt = (double *)calloc(nm, sizeof(double));
  for (int i = 0; i < nat; i++)
    for (int m = 0; m < nm; m++) {
      double s = 0;
      for (int o = 0; o < 4; o++) s += mo[(size_t)(4 * i + o) * nm + m];
      ma[(size_t)i * nm + m] = s;
    }
  for (int i = 0; i < nat; i++)
    for (int m = 0; m < nm; m++) mt[m] += ma[(size_t)i * nm + m];
  double *g = (double *)malloc(sizeof(double) * M);
  double *c = (double *)malloc(sizeof(double) * M);
  double *w = (double *)malloc(sizeof(double) * M);
  orc_gmM(M, g);
  double kTs = p->kT / a;
  double mu = bisect_mu(p, kTs, g, mt);
  coeffs_at(p, mu, kTs, c);
  kpe_weights(M, c, g, w);
  /* populations: computeDiagonalElements (exact mode) + orbital2Atom */
  for (int i = 0; i < nat; i++) {
    double s = 0;
    for (int m = 0; m < M; m++) s += w[m] * ma[(size_t)i * nm + m];
    pop[i] = s;
  }
  double ns = orc_count(M, w, mt);
  /* Tr(P H~): T_m(x) x = (T_{m+1}(x) + T_{|m-1|}(x)) / 2 */
  double tph = w[0] * mt[1];
  for (int m = 1; m < M; m++) tph += w[m] * 0.5 * (mt[m + 1] + mt[m - 1]);
  if (res) {
    res->mu = mu;
    res->mu_ha = a * mu + b;
    res->a = a;
    res->b = b;
    res->emin = emin;
    res->emax = emax;
    res->nelec = 2.0 * ns;
    res->eband = 2.0 * (a * tph + b * ns);
  }
  free(crow);
  free(ccol);
  free(cval);
  free(mo);
  if (!mom_atom) free(ma);
  free(mt);
  free(g);
  free(c);
  free(w);
  return ORC_OK;
}

/* SCC: constructdeltaH deltaH:1414-1487 (dq, Gamma.dq, diagonal shift),
 * driver loop deltaH:1815-1840; updateHtotal's Hamiltonian mixing
 * (mixRatio = 0, :56) is replaced by linear mixing of the populations. */
int orc_scc(int nat, const int *rowptr, const int *col, const double *val, const int *grow,
            const int *gcol, const double *gval, const double *p0, const orc_kpe_params *p,
            double mix, double tol, int maxit, double *pio, int *iters, double *hist,
            orc_kpe_result *res, int nthreads) {
  double *pin = (double *)malloc(sizeof(double) * nat);
  double *pout = (double *)malloc(sizeof(double) * nat);
  double *v = (double *)malloc(sizeof(double) * nat);
  double *shift = (double *)malloc(sizeof(double) * 4 * nat);
  memcpy(pin, pio, sizeof(double) * nat);
  int status = ORC_NOT_CONVERGED, it;
  for (it = 0; it < maxit; it++) {
    for (int i = 0; i < nat; i++) {
      double s = 0;
      for (int k = grow[i]; k < grow[i + 1]; k++) s += gval[k] * (pin[gcol[k]] - p0[gcol[k]]);
      v[i] = s;
      for (int o = 0; o < 4; o++) shift[4 * i + o] = s;
    }
    int st = orc_density(nat, rowptr, col, val, shift, p, pout, NULL, res, nthreads);
    if (st) {
      status = st;
      break;
    }
    double dq = 0;
    for (int i = 0; i < nat; i++) {
      double d = 2.0 * fabs(pout[i] - pin[i]);
      if (d > dq) dq = d;
    }
    if (hist) hist[it] = dq;
    if (dq < tol) {
      memcpy(pio, pout, sizeof(double) * nat);
      status = ORC_OK;
      it++;
      break;
    }
    for (int i = 0; i < nat; i++) pin[i] += mix * (pout[i] - pin[i]);
  }
  if (status == ORC_NOT_CONVERGED) memcpy(pio, pout, sizeof(double) * nat);
  if (iters) *iters = it;
  free(pin);
  free(pout);
  free(v);
  free(shift);
  return status;
}

/* ------------------------------------------------------------------ */
/* dense SP2: P_k = (1+alpha) P_{k-1} - alpha P_{k-1}^2 (getSP2 deltaH:1200-
 * 1206) with alpha = -1 if 2 tr > Ne else +1 (computeAlphas :1277-1281).  */
int orc_sp2_dense(int n, double *X, double nocc, int maxit, double idem_tol, double *idem_out) {
  double *X2 = (double *)malloc(sizeof(double) * (size_t)n * n);
  int it;
  double idem = INFINITY;
  for (it = 0; it < maxit; it++) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      double *row = X2 + (size_t)i * n;
      for (int j = 0; j < n; j++) row[j] = 0;
      for (int k = 0; k < n; k++) {
        double xik = X[(size_t)i * n + k];
        const double *xk = X + (size_t)k * n;
        for (int j = 0; j < n; j++) row[j] += xik * xk[j];
      }
    }
    double tr = 0, tr2 = 0;
    for (int i = 0; i < n; i++) {
      tr += X[(size_t)i * n + i];
      tr2 += X2[(size_t)i * n + i];
    }
    idem = fabs(tr - tr2);
    if (idem < idem_tol) break;
    if (tr > nocc) {
      memcpy(X, X2, sizeof(double) * (size_t)n * n);
    } else {
      for (size_t k = 0; k < (size_t)n * n; k++) X[k] = 2.0 * X[k] - X2[k];
    }
  }
  if (idem_out) *idem_out = idem;
  free(X2);
  return it;
}
