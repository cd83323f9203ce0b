/*
 * ltb.h -- C ABI of the MI355X density-matrix path (KPE / SP2 / SCC).
 *
 * Plain C types only.  Host arrays are caller-owned and copied; device
 * buffers belong to an ltb_ctx.  One ltb_ctx = one HIP stream; calls on one
 * context are not thread-safe (the reference's entry points are collective
 * over PETSC_COMM_WORLD and run in lockstep, which this mirrors).  Every
 * function returns an ltb_status; ltb_last_error() gives the message.
 *
 * Each entry point names the reference function it replaces
 * (pghale/largeTB, withSP2_deltaH.cpp unless stated).
 */
#ifndef LTB_H
#define LTB_H

#ifdef __cplusplus
extern "C" {
#endif

enum ltb_status {
  LTB_OK = 0,
  LTB_ERR_ARG = 1,        /* bad argument / size                         */
  LTB_ERR_GEOMETRY = 2,   /* Si-O closer than 0.01 Bohr, unknown pair (:219-222, :338-341) */
  LTB_ERR_NEIGHBORS = 3,  /* more than maxN_A = 20 neighbours (:357-359) */
  LTB_ERR_NONFINITE = 4,  /* NaN/Inf met in H or a reduction           */
  LTB_NOT_CONVERGED = 5,  /* SCC hit its iteration cap (:1820)           */
  LTB_ERR_DEVICE = 6,     /* HIP runtime error                           */
  LTB_ERR_STATE = 7,      /* call order (e.g. no system loaded)          */
  LTB_ERR_NOMEM = 8
};

enum ltb_lattice {
  LTB_LATTICE_CONFIG = 0,    /* a=4.913 A, c=5.405 A, 3-D periodic           */
  LTB_LATTICE_REFERENCE = 1  /* a=9.725, c=10.7 Bohr, x/y images (:380,:400) */
};

enum ltb_precision {
  LTB_FP64 = 0,       /* recursion vectors and H~ in FP64                    */
  LTB_FP32_ACC64 = 1  /* FP32 vectors and H~, moments accumulated in FP64    */
};

typedef struct ltb_system ltb_system; /* host: geometry, neighbours, H0 (BSR 4x4), Gamma */
typedef struct ltb_ctx ltb_ctx;       /* device context                                    */

typedef struct {
  int L, M, N;             /* cells along R1, R2, R3                           */
  int lattice;             /* enum ltb_lattice                                  */
  double sigma_ang;        /* Gaussian displacement per Cartesian component (A) */
  unsigned long long seed; /* SplitMix64 seed of the displacements              */
} ltb_quartz_desc;

/* ---- host side: structure, neighbour list, H0 and Gamma assembly ------------ */

/* arrangeCoordinates + initializeCoordinates + getHaloAtoms + constructH +
 * initializeForSCF (:669-738, :504-667, :372-501, :969-1066, :854-967). */
int ltb_system_quartz(const ltb_quartz_desc* desc, ltb_system** out);

/* Same assembly from explicit coordinates (Bohr); Z is 14 (Si) or 8 (O).
 * cell9 rows are the periodic vectors; periodic_mask bit d enables images
 * along row d. */
int ltb_system_create(int natoms, const int* Z, const double* xyz_bohr, const double* cell9,
                      int periodic_mask, ltb_system** out);
void ltb_system_destroy(ltb_system* sys);
const char* ltb_system_last_error(void); /* thread-local message of the last failure */

/* sizes: atoms, BSR blocks of H0, nonzeros of Gamma, valence electrons */
int ltb_system_sizes(const ltb_system* sys, int* natoms, int* nnzb, int* gamma_nnz,
                     double* nelectrons);
int ltb_system_coords(const ltb_system* sys, int* Z, double* xyz_bohr);
/* H0 as BSR 4x4, row-major blocks; per row the diagonal block first, then
 * neighbour blocks by ascending column (MatSetValues targets, :993, :1022). */
int ltb_system_hamiltonian(const ltb_system* sys, int* rowptr, int* col, double* val16);
/* Gamma as CSR with the same row order (Gamma_ab, :887-896). */
int ltb_system_gamma(const ltb_system* sys, int* rowptr, int* col, double* val);
/* neutral per-spin populations (Si 2, O 3; arrangeCoordinates :714-720 with
 * the O branch the reference lacks). */
int ltb_system_neutral(const ltb_system* sys, double* p0);

/* ---- MATLAB ASCII files (the reference's dump format) ---------------------------
 * The layout of PETSc's PETSC_VIEWER_ASCII_MATLAB viewer, in which the
 * reference writes GAMMA_AB.m / orbit2Atom.m (:950-963), deltaH.m, netCharge.m,
 * GammaQ_vector.m (:1465-1484) and H.m, Htotal_scc.m, Htotal0.m (:1504-1523):
 * 1-based "i j  v" rows of the scalar entries (row-major, ascending columns),
 * then NAME = spconvert(zzz); vectors one value per row.  Values as %.16e.
 * A BSR matrix with block size bs is written entry by entry (bs = 1: CSR);
 * nrows / ncols count block rows / block columns.  name: a MATLAB identifier. */
int ltb_write_matlab_bsr(const char* path, const char* name, int nrows, int ncols, int bs, const int* rowptr,
                         const int* col, const double* val);
int ltb_write_matlab_vec(const char* path, const char* name, int n, const double* v);
/* Readers for the same files (also the reference's own dumps).  *nnz / *n:
 * capacity in, count out; with NULL outputs they only report sizes.  Indices
 * come back 0-based. */
int ltb_read_matlab_mat(const char* path, long long* nrows, long long* ncols, long long* nnz, int* rows,
                        int* cols, double* vals);
int ltb_read_matlab_vec(const char* path, long long* n, double* v);
const char* ltb_io_last_error(void); /* thread-local message of the last file error */

/* ---- device context ---------------------------------------------------------- */
int ltb_create(int device, ltb_ctx** out);
void ltb_destroy(ltb_ctx* ctx);
const char* ltb_last_error(const ltb_ctx* ctx);

/* Upload H0, Gamma and the orbital->atom map (initializeForH + MatAssembly). */
int ltb_load_system(ltb_ctx* ctx, const ltb_system* sys);

/* The same system assembled on the device from coordinates (Bohr; Z 14 Si /
 * 8 O; cell9 rows = periodic vectors, may be NULL without periodicity):
 * O(N) cell-list neighbour search, 4x4 Slater-Koster blocks, Gamma and the
 * Gershgorin sums in device kernels.  Replaces getHaloAtoms (:372-501, an
 * O(N^2) host scan), get4by4HmatrixElements (:211-349), constructH
 * (:969-1066) and initializeForSCF's Gamma (:854-900); the blocks, pattern
 * and order equal ltb_system_create's.  Periodic rows must be wider than
 * 8 Bohr (one image per pair).  seconds (optional): wall time of the assembly. */
int ltb_load_coords(ltb_ctx* ctx, int natoms, const int* Z, const double* xyz_bohr, const double* cell9,
                    int periodic_mask, double* seconds);
/* sizes and a host copy of the loaded system's H0 (BSR, as
 * ltb_system_hamiltonian) and Gamma (CSR, as ltb_system_gamma) */
int ltb_get_sizes(ltb_ctx* ctx, int* natoms, int* nnzb, int* gamma_nnz, double* nelectrons);
int ltb_get_hamiltonian(ltb_ctx* ctx, int* rowptr, int* col, double* val16);
int ltb_get_gamma(ltb_ctx* ctx, int* rowptr, int* col, double* val);

/* constructdeltaH (:1414-1487): v = Gamma (p - p0) on device (FP64), H_total =
 * H0 + diag(v repeated over the 4 orbitals).  p: per-spin populations [natoms];
 * NULL keeps the populations the context holds (neutral after loading a
 * system, the last mixed populations after an SCC) and recomputes v. */
int ltb_set_populations(ltb_ctx* ctx, const double* p);
/* replace the neutral per-spin populations p0 of Delta q = p - p0 (default
 * Si 2, O 3); the reference's literal values (Si 2, O 0: :714-720) for
 * comparisons with its runs */
int ltb_set_neutral(ltb_ctx* ctx, const double* p0);
/* replace Gamma_AA (per atom) of the loaded system.  The reference sets
 * Gamma_AA = -Us with Us = 1.0 for every atom (initializeForSCF :876, :887), the
 * default here; its Slater-Koster routine declares Hubbard parameters it never
 * uses, Ud = 0.247609 for Si (:261) and Up = 0.4954 for O (:303), which with
 * Gamma_AA = +U give the on-site Coulomb term of standard SCC tight binding. */
int ltb_set_gamma_onsite(ltb_ctx* ctx, const double* gamma_aa);
int ltb_get_shift(ltb_ctx* ctx, double* v_atom);

/* GershGorinParallel (:144-209) on H_total, FP64 device reduction. */
int ltb_gershgorin(ltb_ctx* ctx, double* emin, double* emax);

/* runKPM rescale (:1554-1560): H~ = (H_total - b)/a, stored on device in
 * FP64 and FP32. */
int ltb_rescale(ltb_ctx* ctx, double a, double b);

/* Exact-trace Chebyshev moments of H~ (getKPE recursion :1139-1168 over all
 * unit vectors, as in the commented exact mode :1361-1371):
 *   atom_moments[m*natoms + A] = sum_{j in A} [T_m(H~)]_jj,  m = 0..M
 *   total_moments[m]           = Tr T_m(H~)
 * Either output may be NULL (results stay on device for ltb_density). */
int ltb_kpe_moments(ltb_ctx* ctx, int M, int precision, double* atom_moments,
                    double* total_moments);

typedef struct {
  int M;          /* moments used by P (m < M)                               */
  double kT;      /* Hartree; <= 0 selects the reference step function      */
  int K;          /* quadrature nodes for the Fermi-Dirac coefficients       */
  double Ne;      /* electrons, both spins                                   */
} ltb_fermi_params;

typedef struct {
  double mu;      /* scaled chemical potential (H~ frame)                    */
  double mu_ha;   /* Hartree                                                 */
  double a, b;    /* rescale used                                            */
  double nelec;   /* electrons reproduced by the bisected mu                 */
  double eband;   /* 2 Tr(P H_total), Hartree                                */
} ltb_density_result;

/* computeDensityMatrix (:1401-1411) in exact mode: mu bisected on device in
 * FP64 to Ne (replaces runKPM's DOS search :1663-1685), Fermi-Dirac (or step)
 * coefficients times the gmMs kernel (:1100-1132), per-atom populations
 * (computeDiagonalElements + orbital2Atom, :1296-1399) and band energy, from
 * the moments of the last ltb_kpe_moments call.  pop: per-spin [natoms]. */
int ltb_density(ltb_ctx* ctx, const ltb_fermi_params* fp, double* pop, ltb_density_result* res);

/* diag(P) per orbital (the reference's diagP, computeDiagonalElements exact
 * mode :1361-1379) at the mu of the last ltb_density call.  Needs per-orbital
 * moments: ltb_set_kpe_orbital_moments(ctx, 1) before ltb_kpe_moments. */
int ltb_set_kpe_orbital_moments(ltb_ctx* ctx, int on);
/* Light cones (default on; no counterpart in the reference, whose getKPE
 * recursion :1148-1168 always runs over every row): with more than one
 * unit-vector chunk under the sweep schedule, recursion step n >= 2 of a chunk
 * visits only the atoms within n + 1 hops of the chunk's own atoms in H~'s
 * block graph (a cyclic atom range per chunk and step), since every other row
 * of t_{n+1} is zero.  Moments equal the full sweep's up to the order of the
 * FP64 dot sums; ltb_kpe_stats counts the bytes of the visited rows only. */
int ltb_set_kpe_light_cone(ltb_ctx* ctx, int on);
/* the weights w[m] = c_m(mu) g_m (w[0] = c_0/2) of the last ltb_density call,
 * as getKPE applies them (:1134-1166) -- input of ltb_kpe_apply */
int ltb_kpe_weights(ltb_ctx* ctx, int M, double* w);
int ltb_density_diag(ltb_ctx* ctx, double* diagP);
/* the per-orbital moments themselves: out[m*(j1-j0) + j - j0] = [T_m(H~)]_jj,
 * m = 0..M, j0 <= j < j1, of the last ltb_kpe_moments call (with M, and with
 * per-orbital moments on) -- the diagonal element getKPE's exact mode reads
 * per unit vector (:1361-1371) before the weights are applied */
int ltb_get_orbital_moments(ltb_ctx* ctx, int M, int j0, int j1, double* out);
/* the weights w[m] (as ltb_kpe_weights) at a given chemical potential mu_ha
 * (Hartree; e.g. runKPM's mu from the DOS, :1681-1712) in the frame of the
 * last ltb_rescale: Fermi-Dirac at kT (Hartree) by K-node quadrature, or the
 * reference's step function (:1100-1103) for kT <= 0 */
int ltb_kpe_weights_at(ltb_ctx* ctx, int M, double mu_ha, double kT, int K, double* w);

typedef struct {
  int M;              /* Chebyshev moments                                       */
  double kT;          /* Hartree (config: 0.05 eV)                               */
  int K;              /* FD quadrature nodes                                     */
  double mix;         /* linear charge mixing (config 0.2)                      */
  double tol;         /* |q_out - q_in|_inf in e (config 1e-6)                  */
  int max_iter;       /* cap (reference: 20, :1820)                             */
  int precision;      /* enum ltb_precision for the bulk iterations            */
  double switch_tol;  /* FP32 bulk -> FP64 once |dq| < switch_tol; convergence is
                         then declared on an FP64 iteration only (0: FP32 to the end) */
  int use_graph;      /* capture the device part of an iteration in a hipGraph */
  int verbose;        /* 1: one line per iteration on stderr                     */
  /* Two-level SCC (0: off; ltb_scc_run only).  Every iteration is the same
   * linear mixing step, but most evaluate the density cheaply: the exact-trace
   * populations at order M_lo (< M), plus the defect F_M(q_ref) - F_M_lo(q_ref)
   * taken at the last full-order iteration q_ref.  A full-order iteration
   * comes first and again whenever the corrected map's own |dq| falls below
   * lo_tol (0: tol / 10) or stops falling (its smallest value not 10% lower
   * for 4 iterations: the FP32 rounding level, ~1.5e-7 e at 15,552 atoms),
   * or after 400 low-order iterations; the run converges only on a full-order iteration in
   * the final precision, so it stops at the plain solve's fixed point.  The
   * full-order iterations run in `precision` until one has |dq| < switch_tol,
   * then in FP64; the low-order ones always in `precision`. */
  int M_lo;
  double lo_tol;
} ltb_scc_params;

typedef struct {
  int iterations;
  int fp64_iterations;
  double dq;          /* last |q_out - q_in|_inf (e)                            */
  double mu_ha;
  double eband;       /* Hartree, 2 Tr(P H) of the last iteration (H from q_in)  */
  double eband_fp;    /* Hartree, eband + 2 pop.Gamma(p_out - p_in): E_band at the
                         fixed point to first order in |q_out - q*|             */
  int graph_launches; /* iterations replayed from a captured hipGraph            */
  double seconds;     /* wall time of the loop                                   */
  int lo_iterations;  /* two-level SCC: iterations on the corrected order-M_lo
                         density (included in iterations)                        */
} ltb_scc_result;

/* SCC driver (main :1815-1840): loop of constructdeltaH, runKPM (bounds +
 * rescale), computeDensityMatrix and linear mixing of the populations (in
 * place of updateHtotal's Hamiltonian mixing, :1489-1526).  p_guess (per
 * spin) may be NULL (neutral start).  p_out receives the final q_out / 2.
 * hist (optional, [max_iter]) receives |dq| per iteration. */
int ltb_scc_run(ltb_ctx* ctx, const ltb_scc_params* prm, const double* p_guess, double* p_out,
                double* hist, ltb_scc_result* res);
/* The per-spin populations the next SCC iteration would start from: q_in / 2
 * after the last mixing step of ltb_scc_run (the reference's mixed charges,
 * updateHtotal :1489-1526 in charge form).  Passing them as p_guess to a new
 * ltb_scc_run continues the same iteration sequence exactly (linear mixing
 * keeps no other state), so a long SCC can run in segments. */
int ltb_get_mixed_populations(ltb_ctx* ctx, double* p);

/* ---- many independent structures in one SCC ----------------------------------
 * The reference solves one structure per job (main :1742-1849, testjob2.pbs);
 * a batch runs the same per-structure SCC as ltb_scc_run for every structure
 * at once: its own Gershgorin bounds and rescale, mu, Gamma shift, mixing,
 * FP32 -> FP64 switch and convergence, while every Chebyshev recursion step
 * of all unconverged structures is one kernel launch (plus one for those in
 * FP64).  Converged structures drop out of the launches. */
typedef struct ltb_batch ltb_batch;

/* concatenates the structures (each keeps its own Gamma and electron count)
 * on ctx's device; the systems may be destroyed afterwards */
int ltb_batch_create(ltb_ctx* ctx, int nstruct, const ltb_system* const* systems, ltb_batch** out);
void ltb_batch_destroy(ltb_batch* b);
/* total atoms (sum over structures) and per-structure atom offsets [nstruct+1] */
int ltb_batch_sizes(const ltb_batch* b, int* nstruct, int* natoms_total, int* atom_offsets);
/* p_guess / p_out: per-spin populations of all atoms, concatenated in
 * structure order [natoms_total] (p_guess NULL = neutral).  Per structure
 * [nstruct] (each optional): iterations, eband_fp (Hartree, as ltb_scc_result),
 * mu_ha and status (LTB_OK, LTB_NOT_CONVERGED, LTB_ERR_NONFINITE).  total:
 * iterations of the whole batch, those with an FP64 launch, graph replays,
 * wall time.  Returns LTB_NOT_CONVERGED if any structure hit max_iter.
 * Two-level (prm->M_lo > 0): a structure recursing in FP32 runs the two-level
 * SCC of ltb_scc_run (low-order iterations at M_lo corrected by the defect of
 * its last full-order iteration, whose order-M_lo populations come from the
 * first M_lo + 1 moments of the same recursion) until a full-order iteration
 * has |dq| < switch_tol, then plain FP64 iterations; the recursion of a
 * batch iteration stops after step M_lo/2 for the structures in a low-order
 * iteration.  total->lo_iterations counts batch iterations with only
 * low-order (or no) recursions past step M_lo/2. */
int ltb_batch_scc_run(ltb_batch* b, const ltb_scc_params* prm, const double* p_guess, double* p_out,
                      int* iterations, double* eband_fp, double* mu_ha, int* status,
                      ltb_scc_result* total);

/* step-kernel launches, their summed duration (s, HIP events on the context
 * stream) and algorithmic bytes in the eager iterations of the last
 * ltb_batch_scc_run while ltb_set_kpe_timing is on (graph replays are not
 * timed).  Bytes per structure and step: the valid columns of the vectors
 * read by the gathers, read as t_{n-1} and written as t_{n+1} (1, 2, 3
 * streams for steps 0, 1, >1) plus its H~ blocks, columns and row pointers. */
int ltb_batch_stats(const ltb_batch* b, long long* launches, double* seconds, double* bytes);

/* ---- stochastic-trace mode (the reference's default estimator) ----------------
 * Moments r^T T_m(H~) r, m = 0..M, of given vectors R ([nvec][norb], host) by
 * the exact-trace block recursion with explicit start vectors (moment
 * doubling): moments [nvec][M+1] and trace[m] = sum over the vectors (either
 * may be NULL).  With R = unit vectors this is the exact trace; with Gaussian
 * R it is runKPM's momentsTrace (:1607-1633).  H~ from ltb_rescale. */
int ltb_kpe_block_moments(ltb_ctx* ctx, int M, int nvec, const double* R, int precision, double* moments,
                          double* trace);
/* nvec x n standard normal numbers (vector-major) from the SplitMix64 ->
 * Box-Muller stream of seed -- the vectors the two calls below use */
int ltb_random_vectors(unsigned long long seed, int nvec, int n, double* R);
/* runKPM's stochastic moments (:1607-1633) with nvec seeded Gaussian vectors,
 * drawn on the device (the ltb_random_vectors stream up to last-ulp libm
 * differences) */
int ltb_kpe_stochastic(ltb_ctx* ctx, int M, int nvec, unsigned long long seed, int precision, double* trace);
/* computeDiagonalElements, random mode (:1338-1392): diagP = sum_r r.*(P r) ./
 * sum_r r.*r, P r = sum_m w[m] T_m(H~) r (getKPE on the block, w as for
 * ltb_kpe_apply), the same seeded vectors */
int ltb_density_stochastic(ltb_ctx* ctx, int M, const double* w, int nvec, unsigned long long seed,
                           int precision, double* diagP);
/* computeAlphas' random-mode trace (:1240-1275) at SP2 depth 0:
 * trace = sum_r r^T (P r) / nvec over the same seeded vectors (the Hutchinson
 * estimate of Tr P), P r = sum_m w[m] T_m(H~) r */
int ltb_trace_stochastic(ltb_ctx* ctx, int M, const double* w, int nvec, unsigned long long seed, int precision,
                         double* trace);
/* runKPM's DOS and mu (:1634-1712), host arithmetic on nmom trace moments:
 * Jackson kernel g_n (N = nmom), DOS at x_k = cos(pi (k+1/2)/ntilde), energies
 * a x_k + b; mu where the DOS area from the bottom reaches Ne/(2 nsize) of the
 * total; emin / emax where DOS/nsize first reaches 0.001 from either end (the
 * reference's bound tightening).  Any output may be NULL. */
int ltb_kpm_dos(const double* moments, int nmom, int ntilde, double a, double b, int nsize, double nelectrons,
                double* energies, double* dos, double* mu, double* emin, double* emax);
/* dosfile.txt rows "E, DOS" (:1714-1717) */
int ltb_write_dos(const char* path, int n, const double* energies, const double* dos);

/* getKPE (:1068-1185) on a block of vectors: Y = sum_m w[m] T_m(H~) X, with
 * w = (c_0/2, c_1 g_1, ...).  X, Y: [ncols][norb] column vectors (host). */
int ltb_kpe_apply(ltb_ctx* ctx, int M, const double* w, int ncols, const double* X, double* Y,
                  int precision);
/* Allocation-free form of ltb_kpe_apply.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under graph capture,
 * serving loops, harnesses that forbid hipMalloc).  X and Y are DEVICE
 * buffers of the caller, [ncols][norb] FP64 column vectors; w stays a host
 * array.  The recursion vectors and theta live in one fixed 96 MiB scratch
 * array per GPU (norb * ld * (2 * sizeof(T) + 8) bytes, ld = ncols rounded up
 * to 64, T the recursion precision); larger calls fail with LTB_ERR_ARG and a
 * message naming the largest column count that fits.  It calls no hipMalloc
 * or hipFree.  Calls on one GPU share the array, so calls that overlap in
 * time on one GPU (other contexts, other threads) race; each call clears what
 * it reads before reading it.  Returns after the context stream has drained. */
int ltb_kpe_apply_noalloc(ltb_ctx* ctx, int M, const double* w, int ncols, const double* X, double* Y,
                          int precision);

/* SP2 (getSP2 :1187-1215 / computeAlphas :1217-1294) on an explicit dense
 * matrix: X0 = (emax I - H_total)/(emax - emin) from the Gershgorin bounds,
 * X <- X^2 if 2 tr X > Ne else 2X - X^2, FP32 MFMA GEMMs until the
 * idempotency error stops falling (or max_iter), then fp64_tail FP64 MFMA
 * iterations.  The FP64 tail restores idempotency and the trace; the occupied
 * subspace keeps the FP32 rounding (~1e-6 per diagonal element).  max_iter = 0
 * runs FP64 from the start, with fp64_tail as its cap.  diagP (optional, [norb]) and pop (optional, per-spin [natoms])
 * receive diag(X) and its per-atom sums. */
typedef struct {
  int iterations, fp64_iterations;
  double trace;   /* tr X at the end (per spin)                  */
  double idem;    /* |tr X - tr X^2| of the final iterate          */
  double gemm_seconds; /* time inside the X^2 products          */
  double gemm_flops;   /* MFMA flops executed, summed: 2 x 128^2 x n per formed
                          128 x 128 tile, n = norb rounded up to 128; the tiles on
                          and above the diagonal are formed (X^2 is symmetric) */
} ltb_sp2_result;
int ltb_sp2_dense(ltb_ctx* ctx, double Ne, int max_iter, int fp64_tail, double* diagP, double* pop,
                  ltb_sp2_result* res);

/* Timing of the last exact-trace moment computation made while timing was on
 * (an ltb_kpe_moments call, or the first iteration of an ltb_scc_run):
 * number of recursion-step launches, their summed device time (HIP events on
 * the context stream), and the algorithmic bytes they move (3 vector streams
 * + H~ blocks + row pointers, over the rows visited). */
int ltb_kpe_stats(ltb_ctx* ctx, int* launches, double* step_seconds, double* alg_bytes);

/* Record HIP events around every recursion-step launch (on the context
 * stream) so ltb_kpe_stats reports their summed device time.  ltb_scc_run
 * records them in its first iteration only (events around every launch
 * lengthen an iteration by ~1.4%). */
int ltb_set_kpe_timing(ltb_ctx* ctx, int on);

/* Work mapping of the Chebyshev block-SpMM kernel.  No reference counterpart
 * (the reference runs one vector at a time through PETSc's MatMult). */
enum ltb_kpe_schedule {
  LTB_SCHED_AUTO = 0,
  LTB_SCHED_BLOCKED = 1, /* workgroup = (column group, consecutive atoms)          */
  LTB_SCHED_SWEEP = 2    /* resident grid; each XCD sweeps its column groups       */
};
typedef struct {
  int chunk_cols;     /* unit-vector columns in flight per chunk (0 = auto)      */
  int cols_per_lane;  /* V in {1, 2, 4}: a wave covers 64*V columns (0 = auto)   */
  int schedule;       /* enum ltb_kpe_schedule                                   */
  int blocks_per_cu;  /* sweep: resident workgroups per CU (0 = auto)            */
  int atoms_per_wave; /* blocked: consecutive atoms per wave (0 = auto)          */
  int nontemporal;    /* 1: t_{n-1}, t_{n+1} bypass the caches; 0: cached; -1 auto */
  int order;          /* sweep visiting order of the atoms: 0 auto (= 2), 1 spatial
                         strips (bonded atoms close in sequence), 2 atom index   */
  int lo_chunk_cols;  /* chunk width of the two-level SCC's low-order recursions
                         (order M_lo), 0 = chunk_cols.  Under light cones a short
                         recursion moves fewer bytes in narrower chunks; set by
                         ltb_kpe_autotune.  Batches ignore it.                  */
} ltb_kpe_tiling;
/* The context keeps one tiling per recursion precision (FP32, FP64); this
 * sets both. */
int ltb_set_kpe_tiling(ltb_ctx* ctx, const ltb_kpe_tiling* t);
/* the tiling the kernel would use for the loaded system (auto fields resolved) */
int ltb_get_kpe_tiling(ltb_ctx* ctx, int precision, ltb_kpe_tiling* out);
/* Time `trial_steps` recursion steps of every candidate tiling on the loaded
 * system, keep the fastest as the tiling of `precision` (the other
 * precision's tiling is left as it was) and report it with its algorithmic
 * GB/s.  With light cones on and a sweep over several chunks it then times a
 * whole order-32 recursion at the chosen chunk width and at half of it and
 * keeps the faster as lo_chunk_cols.  The trials overwrite the recursion
 * buffers and the stored moments: run ltb_kpe_moments again before
 * ltb_density. */
int ltb_kpe_autotune(ltb_ctx* ctx, int precision, int trial_steps, ltb_kpe_tiling* best,
                     double* best_gbs);

/* Tiling of a batch's step kernel.  Column groups are fixed (256 columns, 4
 * per lane in FP32 and 2 x 2 in FP64, then one 64-, 128- or 256-wide group
 * for the columns left);
 * atoms_per_wave, blocks_per_cu and nontemporal are taken (0 / -1 keep the
 * current value).  No reference counterpart: the reference has one structure
 * per job and no tiling. */
int ltb_batch_set_tiling(ltb_batch* b, const ltb_kpe_tiling* t);
int ltb_batch_get_tiling(const ltb_batch* b, ltb_kpe_tiling* out);
/* Time `trial_steps` steps of every candidate (atoms_per_wave 4/8/16 x
 * blocks_per_cu 2/4/8 x nontemporal 0/1) over all structures at their neutral
 * start, keep the fastest and report its algorithmic GB/s. */
int ltb_batch_autotune(ltb_batch* b, int precision, int trial_steps, ltb_kpe_tiling* best,
                       double* best_gbs);

/* ---- learned SCC starting guess ----------------------------------------------
 * No reference counterpart: the reference starts every SCC from the neutral
 * charges (arrangeCoordinates :714-720).  A per-species MLP on local
 * descriptors predicts converged charges; the plain neutral start remains the
 * fallback and the answer the guessed runs are checked against. */
#define LTB_MLP_NFEAT 80 /* 2 x 32 radial Gaussians (per neighbour species) + 16 angular */
typedef struct ltb_mlp ltb_mlp;
int ltb_mlp_num_params(void);
int ltb_mlp_create(ltb_ctx* ctx, unsigned long long seed, ltb_mlp** out);
void ltb_mlp_destroy(ltb_mlp* mlp);
int ltb_mlp_get_params(ltb_mlp* mlp, float* params);
int ltb_mlp_set_params(ltb_mlp* mlp, const float* params); /* also resets the Adam state */
/* descriptors of every atom of sys (cutoff 6 A, periodic images), [natoms][80] */
int ltb_descriptors(ltb_ctx* ctx, const ltb_system* sys, float* feat);
/* training / evaluation set: atoms grouped in nseg structures (seg[0..nseg]),
 * qtot[s] = electrons of structure s, target = charges in e (or NULL) */
int ltb_mlp_set_data(ltb_mlp* mlp, int natoms, const float* feat, const int* Z, int nseg,
                     const int* seg, const double* qtot, const double* target);
int ltb_mlp_fit_normalization(ltb_mlp* mlp); /* feature mean/sd and charge shift per species */
int ltb_mlp_forward(ltb_mlp* mlp, double* q);  /* charges (e), neutral per structure */
int ltb_mlp_loss_grad(ltb_mlp* mlp, double* loss, float* grad /* [trainable] or NULL */);
int ltb_mlp_adam(ltb_mlp* mlp, int steps, float lr, double* loss_hist /* [steps] or NULL */);
/* per-spin population guess for ltb_scc_run (replaces the data set) */
int ltb_mlp_predict(ltb_mlp* mlp, const ltb_system* sys, double* p_guess);

#ifdef __cplusplus
}
#endif
#endif
