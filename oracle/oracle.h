/*
 * oracle.h -- CPU FP64 restatement of pghale/largeTB's density-matrix path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  Nothing in largetb_amd/ links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pins (what checks
 * each function against something other than itself) are listed in the header
 * comment of oracle.c.
 */
#ifndef LTB_ORACLE_H
#define LTB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXN 20 /* maxN_A, withSP2_deltaH.cpp:26 */

/* status codes (same meaning as the product's) */
#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_GEOMETRY 2
#define ORC_ERR_NEIGHBORS 3
#define ORC_ERR_NONFINITE 4
#define ORC_NOT_CONVERGED 5

/* geometry modes */
#define ORC_MODE_CONFIG 0    /* a=4.913 A, c=5.405 A, 3-D periodic, Gaussian noise */
#define ORC_MODE_REFERENCE 1 /* a=9.725, c=10.7 Bohr, x/y images only, no noise   */

uint64_t orc_splitmix64(uint64_t *state);
double orc_normal(uint64_t *state, int *have, double *spare);

/* alpha-quartz supercell: xyz[3*nat] (Bohr), Z[nat] (14 Si / 8 O), cell[9]
 * (rows = supercell vectors, Bohr).  initializeCoordinates deltaH:504-667. */
int orc_quartz(int L, int M, int N, int mode, double sigma_ang, uint64_t seed,
               double *xyz, int *Z, double *cell);

/* brute-force neighbour search over images (getHaloAtoms deltaH:372-501,
 * ModifyNList :355-370).  periodic_mask bit d = images along cell row d.
 * nn[nat], nbr[nat*ORC_MAXN] (atom index), nd[nat*ORC_MAXN*3] (r_j - r_i). */
int orc_neighbors(int nat, const double *xyz, const double *cell, int periodic_mask,
                  int *nn, int *nbr, double *nd);

/* 4x4 Slater-Koster block, row-major (get4by4HmatrixElements deltaH:211-349) */
int orc_sk_block(int zi, int zj, double x, double y, double z, double *h16);

/* H assembly as BSR 4x4 (constructH deltaH:969-1066).  Blocks in row i are
 * ordered: diagonal first, then neighbours in neighbour-list order; a repeated
 * (i,j) pair overwrites the earlier block (PETSc INSERT_VALUES semantics).
 * Returns nnzb (>=0) or -status.  rowptr[nat+1], col[cap], val[cap*16]. */
int orc_build_h(int nat, const int *Z, const int *nn, const int *nbr, const double *nd,
                int cap, int *rowptr, int *col, double *val);

/* Gamma as CSR (initializeForSCF deltaH:854-900): diag -Us (Us=1.0), off-diag
 * 1/r for listed neighbours, INSERT semantics.  Returns nnz or -status. */
int orc_build_gamma(int nat, const int *nn, const int *nbr, const double *nd, double Us,
                    int cap, int *rowptr, int *col, double *val);

/* Gershgorin bounds of H + diag(shift) (GershGorinParallel deltaH:144-209),
 * shift per orbital or NULL. */
int orc_gershgorin(int nat, const int *rowptr, const int *col, const double *val,
                   const double *shift, double *emin, double *emax);

/* Kernels and coefficients */
void orc_gmM(int M, double *g);                  /* getKPE gmMs, deltaH:1105-1132 */
void orc_jackson(int N, double *g);              /* runKPM gns, deltaH:1593-1595 */
void orc_cheb_nodes(int n, double *x);           /* runKPM xk, deltaH:1589-1591 */
void orc_step_coeffs(int M, double mu, double *c); /* mu_mvals, deltaH:1100-1103 */
void orc_fd_coeffs(int M, double mu, double kT, int K, double *c); /* Fermi-Dirac */

/* scalar CSR of (H + diag(shift) - b)/a from the BSR (runKPM rescale
 * deltaH:1554-1560). Arrays sized nat*4+1 and nnzb*16. */
int orc_rescaled_csr(int nat, const int *rowptr, const int *col, const double *val,
                     const double *shift, double a, double b, int *crow, int *ccol, double *cval);

/* Exact-trace Chebyshev moments by the getKPE op sequence (deltaH:1139-1168)
 * applied to unit vectors e_j for j in [j0, j1): MatMult + VecCopy + VecScale +
 * VecAXPY + VecCopy + VecCopy per step; the diagonal element x_n[j] is read
 * where getKPE does its theta VecAXPY.  mom_orb[(j-j0)*(nm)+n] = [T_n]_jj for
 * n < nm.  nthreads<=0 -> OpenMP default. */
int orc_kpe_diag_moments(int n, const int *crow, const int *ccol, const double *cval, int nm,
                         int j0, int j1, double *mom_orb, int nthreads);
/* the same for the listed columns: mom_orb[k*nm + n] = [T_n]_{cols[k] cols[k]} */
int orc_kpe_diag_moments_cols(int n, const int *crow, const int *ccol, const double *cval, int nm,
                              int ncols, const int *cols, double *mom_orb, int nthreads);

/* getKPE (deltaH:1068-1185): out = w0*x + sum_{m>=1} w_m T_m x, with
 * w = (c0/2, c1 g1, ..., c_{M-1} g_{M-1}) given in w[M]. */
int orc_kpe_apply(int n, const int *crow, const int *ccol, const double *cval, int M,
                  const double *w, const double *x, double *out);

/* electron count per spin from total moments mu[M]: sum w_m mu_m */
double orc_count(int M, const double *w, const double *mu_tot);

typedef struct {
  int M;          /* Chebyshev moments (P uses m < M) */
  double kT;      /* Hartree; <=0 -> step function (reference) */
  int K;          /* quadrature nodes for FD coefficients */
  double Ne;      /* electrons (both spins) */
  int bisect_iters;
} orc_kpe_params;

typedef struct {
  double mu;      /* scaled chemical potential */
  double mu_ha;   /* in Hartree */
  double a, b;    /* rescale */
  double emin, emax;
  double nelec;   /* 2*sum w mu achieved */
  double eband;   /* 2 Tr(P H), Hartree */
} orc_kpe_result;

/* full density-matrix step for H + diag(shift): Gershgorin, rescale, exact
 * moments, mu bisection, per-atom populations pop[nat] (per spin, as the
 * reference's chargeOnAtom, deltaH:1385), band energy.  mom_atom (optional,
 * [nat*(M+1)]) receives per-atom moments. */
int orc_density(int nat, const int *rowptr, const int *col, const double *val,
                const double *shift, const orc_kpe_params *p, double *pop, double *mom_atom,
                orc_kpe_result *res, int nthreads);

/* SCC loop with linear charge mixing.  p_in: per-spin populations, in/out
 * (input = starting guess, output = final q_out/2).  p0: neutral per-spin
 * populations.  Gamma CSR.  Returns ORC_OK or ORC_NOT_CONVERGED. hist[maxit]
 * receives |dq|inf (in e) per iteration (optional). */
int orc_scc(int nat, const int *rowptr, const int *col, const double *val, const int *grow,
            const int *gcol, const double *gval, const double *p0, const orc_kpe_params *p,
            double mix, double tol, int maxit, double *p_io, int *iters, double *hist,
            orc_kpe_result *res, int nthreads);

/* dense SP2 (Niklasson 2002, trace-correcting branch rule of computeAlphas
 * deltaH:1277-1281): X (n x n, row-major) in/out. Returns iterations or -status. */
int orc_sp2_dense(int n, double *X, double nocc, int maxit, double idem_tol, double *idem_out);

#ifdef __cplusplus
}
#endif
#endif
