// ltb_petsc.hpp -- the reference's operator API (pghale/largeTB,
// withSP2_deltaH.cpp) restated over the C ABI in ltb.h.
//
// The reference keeps its state in globals (:26-64) and calls PETSc for every
// Mat/Vec operation.  Here one State object holds what those globals hold,
// Vec is a host vector, Mat a host copy of a BSR/CSR matrix, and every entry
// point keeps the reference's name, argument meaning and int return code
// (0 = success, else an ltb_status).  The device work behind each call runs
// through the C ABI; there is no CPU fallback.
#ifndef LTB_PETSC_HPP
#define LTB_PETSC_HPP

#include <string>
#include <vector>

#include "ltb.h"

namespace ltb_petsc {

typedef int PetscErrorCode;
typedef std::vector<double> Vec;

struct Mat {                  // host copy; bs = 4 (BSR, H) or 1 (CSR, Gamma)
  int nrows = 0, bs = 1;
  std::vector<int> rowptr, col;
  std::vector<double> val;
};

struct Settings {
  int L = 1, M = 1, N = 1;              // argv L M N (:1747-1756)
  int lattice = LTB_LATTICE_CONFIG;
  double sigma_ang = 0.0;
  unsigned long long seed = 1;
  int moments = 200;                    // getKPE M (:1094)
  double kT_ev = 0.05;                  // <= 0: the reference's step function (:1100-1103)
  int quad_nodes = 8192;
  int maxsp2 = 1;                       // computeDensityMatrix (:1402)
  double force_alpha = 0.0;             // != 0: every SP2 level uses this alpha instead of computeAlphas' rule
  std::vector<double> force_alphas;     // non-empty: level k uses force_alphas[min(k, size - 1)] (overrides)
  double kpe_scale = 1.0;               // getKPE's output scale: 1 as here (:1068-1185); 2 as withSP2.cpp's
                                        // getKPE, which returns 2 P x, the spin-summed density (SP2:939)
  int precision = LTB_FP64;
  double mix = 0.2;                     // linear charge mixing (replaces mixRatio = 0.0, :56)
  double tol = 1e-6;                    // |dq|inf in e (reference: 0.1, :1839)
  int max_scf = 20;                     // :1820
  int device = 0;
  bool verbose = true;
  int nrandoms = 0;                     // nRandoms (argv[4], :1757): 0 = exact trace over unit vectors;
                                        // > 0 = runKPM's stochastic DOS and mu (:1562-1712) and the
                                        // random-vector diagonal (:1338-1392) with this many vectors
  int dos_moments = 1000;               // Nmoments of the stochastic DOS (:1565)
  bool reference_literal = false;       // the reference's own SCC, for comparison with its runs: neutral
                                        // O charge 0 (:714-720 has no O branch), the loop measuring
                                        // successive outputs |q_k - q_(k-1)| (:1830-1838); mix then plays
                                        // mixRatio (Hamiltonian mixing equals charge mixing, Gamma linear)
  std::string dump_dir;                 // non-empty: write the reference's MATLAB dumps there (H.m,
                                        // GAMMA_AB.m, deltaH.m, netCharge.m, GammaQ_vector.m)
  int two_level_order = 0;              // > 0 (exact trace, maxsp2 1): scf runs the device two-level SCC
                                        // (ltb_scc_run with M_lo = this order; FP32 full-order iterations
                                        // until |dq| < 1e-2 e with --fp32, FP64 after)
};

struct State {
  Settings opt;
  ltb_system* sys = nullptr;
  ltb_ctx* ctx = nullptr;
  int Natoms = 0, nSize = 0;
  double Nelectrons = 0;
  double muval = 0;                     // scaled chemical potential (:43)
  double a = 1, b = 0;                  // H~ = (H - b)/a
  double eband = 0;                     // 2 Tr(P H), Hartree (exact-trace mode)
  double mu_ha = 0;                     // chemical potential, Hartree
  unsigned long long draws = 0;         // random-vector sets drawn so far (each gets its own seed)
  Vec dosEnergies, dosValues;           // runKPM's DOS (stochastic mode)
  Mat H, Gamma_ab;
  Vec neutralCharges, chargeOnAtom, previousCharges, netCharges, gammaq, diagP;
  Vec mixedCharges;                     // populations H_total is built from
  std::vector<double> alphavals;
  std::string last_error;
  // wall seconds of runKPM / computeDensityMatrix (every call returns after the
  // device has finished): total, number of calls, and the last call
  double t_kpm = 0, t_density = 0, t_kpm_last = 0, t_density_last = 0;
  int n_kpm = 0, n_density = 0;
};

PetscErrorCode arrangeCoordinates(State& s);                         // :669-738
PetscErrorCode initializeForH(State& s);                             // :786-833
PetscErrorCode constructH(State& s);                                 // :969-1066
PetscErrorCode initializeForSCF(State& s);                           // :854-967
PetscErrorCode GershGorinParallel(State& s, double* mn, double* mx); // :144-209
PetscErrorCode runKPM(State& s);                                     // :1528-1723
PetscErrorCode getKPE(State& s, const Vec& in, Vec& out);            // :1068-1185
PetscErrorCode getSP2(State& s, const Vec& in, Vec& out, int localsp2max, int currentsp2);  // :1187-1215
PetscErrorCode computeAlphas(State& s);                              // :1217-1294
PetscErrorCode computeDiagonalElements(State& s);                    // :1296-1399
PetscErrorCode computeDensityMatrix(State& s);                       // :1401-1411
PetscErrorCode constructdeltaH(State& s);                            // :1414-1487
PetscErrorCode updateHtotal(State& s);                               // :1489-1526
PetscErrorCode scf(State& s, int* iterations, double* delta);        // main loop :1815-1840
void clearAll(State& s);                                             // :741-784

}  // namespace ltb_petsc
#endif
