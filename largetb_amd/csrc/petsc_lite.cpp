// The reference's entry points (withSP2_deltaH.cpp) over the C ABI.  Each
// function keeps the reference's name, the role of its arguments and its
// integer return code; the Mat/Vec work of each runs on the device.
#include "../../include/ltb_petsc.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace ltb_petsc {
namespace {

int fail(State& s, int code) {
  if (code != LTB_OK) s.last_error = s.ctx ? ltb_last_error(s.ctx) : ltb_system_last_error();
  return code;
}

int io_fail(State& s, int code) {
  if (code != LTB_OK) s.last_error = ltb_io_last_error();
  return code;
}

#define CHK(s, call)                  \
  do {                                \
    int e_ = (call);                  \
    if (e_ != LTB_OK) return fail(s, e_); \
  } while (0)

#define CHKIO(s, call)                   \
  do {                                   \
    int e_ = (call);                     \
    if (e_ != LTB_OK) return io_fail(s, e_); \
  } while (0)

// apply getSP2 at depth `depth` to the unit vectors e_j0..e_j1-1 and return
// their diagonal elements [P_depth]_jj (the exact mode of :1266-1275, :1361-1371)
int diag_block(State& s, int depth, int j0, int j1, std::vector<double>& out) {
  for (int j = j0; j < j1; j++) {
    Vec e(s.nSize, 0.0), y;
    e[j] = 1.0;
    CHK(s, getSP2(s, e, y, depth, depth));
    out[j] = y[j];
  }
  return LTB_OK;
}

std::string dump_path(const State& s, const char* file) { return s.opt.dump_dir + "/" + file; }

// wall time of one phase call, added to the state's totals on every return path
struct PhaseTimer {
  double& total;
  double& last;
  int& calls;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  PhaseTimer(double& t, double& l, int& n) : total(t), last(l), calls(n) {}
  ~PhaseTimer() {
    last = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    total += last;
    calls++;
  }
};

// the reference prints these two after the first density (main :1800-1812)
void print_phase_times(const State& s) {
  if (!s.opt.verbose) return;
  std::printf("time KPM : %g\n", s.t_kpm_last);
  std::printf("time SP2 : %g\n", s.t_density_last);
}

// getKPE's coefficients c_m(mu) g_m (:1094-1132): from the bisected mu of the
// exact mode, or at runKPM's DOS mu in stochastic mode
int kpe_weights(State& s, std::vector<double>& w) {
  w.assign(s.opt.moments, 0.0);
  if (s.opt.nrandoms > 0) {
    const double kT = s.opt.kT_ev > 0 ? s.opt.kT_ev / 27.211386 : 0.0;
    CHK(s, ltb_kpe_weights_at(s.ctx, s.opt.moments, s.mu_ha, kT, s.opt.quad_nodes, w.data()));
  } else {
    CHK(s, ltb_kpe_weights(s.ctx, s.opt.moments, w.data()));
  }
  return LTB_OK;
}

// computeDiagonalElements' random mode (:1338-1392) at SP2 depth `depth`:
// diag = sum_r r.*(P_depth r) ./ sum_r r.*r over opt.nrandoms seeded vectors
int random_diag(State& s, int depth, std::vector<double>& out) {
  const int nR = s.opt.nrandoms, n = s.nSize;
  const unsigned long long seed = s.opt.seed + 1000003ULL * ++s.draws;
  out.assign(n, 0.0);
  if (depth == 0) {
    std::vector<double> w;
    CHK(s, kpe_weights(s, w));
    CHK(s, ltb_density_stochastic(s.ctx, s.opt.moments, w.data(), nR, seed, s.opt.precision, out.data()));
    for (double& x : out) x *= s.opt.kpe_scale;
    return LTB_OK;
  }
  std::vector<double> R((size_t)nR * n), num(n, 0.0), den(n, 0.0);
  if (ltb_random_vectors(seed, nR, n, R.data()) != LTB_OK) return LTB_ERR_ARG;
  for (int k = 0; k < nR; k++) {
    Vec r(R.begin() + (size_t)k * n, R.begin() + (size_t)(k + 1) * n), y;
    CHK(s, getSP2(s, r, y, depth, depth));
    for (int j = 0; j < n; j++) {
      num[j] += r[j] * y[j];
      den[j] += r[j] * r[j];
    }
  }
  for (int j = 0; j < n; j++) out[j] = num[j] / den[j];
  return LTB_OK;
}

// computeAlphas' trace estimate at SP2 depth `depth` (:1240-1275): the mean
// of r^T P_depth r over opt.nrandoms seeded vectors
int random_trace(State& s, int depth, double* tr) {
  const int nR = s.opt.nrandoms, n = s.nSize;
  const unsigned long long seed = s.opt.seed + 1000003ULL * ++s.draws;
  if (depth == 0) {
    std::vector<double> w;
    CHK(s, kpe_weights(s, w));
    CHK(s, ltb_trace_stochastic(s.ctx, s.opt.moments, w.data(), nR, seed, s.opt.precision, tr));
    *tr *= s.opt.kpe_scale;
    return LTB_OK;
  }
  std::vector<double> R((size_t)nR * n);
  if (ltb_random_vectors(seed, nR, n, R.data()) != LTB_OK) return LTB_ERR_ARG;
  double sum = 0;
  for (int k = 0; k < nR; k++) {
    Vec r(R.begin() + (size_t)k * n, R.begin() + (size_t)(k + 1) * n), y;
    CHK(s, getSP2(s, r, y, depth, depth));
    for (int j = 0; j < n; j++) sum += r[j] * y[j];
  }
  *tr = sum / nR;
  return LTB_OK;
}
}  // namespace

PetscErrorCode arrangeCoordinates(State& s) {
  ltb_quartz_desc d{s.opt.L, s.opt.M, s.opt.N, s.opt.lattice, s.opt.sigma_ang, s.opt.seed};
  if (s.sys) ltb_system_destroy(s.sys);
  s.sys = nullptr;
  CHK(s, ltb_system_quartz(&d, &s.sys));
  int nnzb, gnnz;
  CHK(s, ltb_system_sizes(s.sys, &s.Natoms, &nnzb, &gnnz, &s.Nelectrons));
  s.nSize = 4 * s.Natoms;
  s.neutralCharges.assign(s.Natoms, 0.0);
  CHK(s, ltb_system_neutral(s.sys, s.neutralCharges.data()));
  if (s.opt.reference_literal) {  // localAtoms == 8 has no branch at :714-720
    std::vector<int> Z(s.Natoms);
    CHK(s, ltb_system_coords(s.sys, Z.data(), nullptr));
    for (int i = 0; i < s.Natoms; i++)
      if (Z[i] == 8) s.neutralCharges[i] = 0.0;
  }
  s.chargeOnAtom = s.neutralCharges;
  s.mixedCharges = s.neutralCharges;
  s.previousCharges = s.neutralCharges;
  return LTB_OK;
}

PetscErrorCode initializeForH(State& s) {
  if (!s.sys) return LTB_ERR_STATE;
  if (!s.ctx) CHK(s, ltb_create(s.opt.device, &s.ctx));
  CHK(s, ltb_load_system(s.ctx, s.sys));
  if (s.opt.reference_literal) CHK(s, ltb_set_neutral(s.ctx, s.neutralCharges.data()));
  CHK(s, ltb_set_kpe_orbital_moments(s.ctx, 1));
  return LTB_OK;
}

PetscErrorCode constructH(State& s) {
  // the blocks are assembled with the structure (ltb_system_quartz); this is
  // the host copy the reference's MatView dumps would show (:1504-1509)
  int nnzb, gnnz;
  double ne;
  CHK(s, ltb_system_sizes(s.sys, &s.Natoms, &nnzb, &gnnz, &ne));
  s.H.nrows = s.Natoms;
  s.H.bs = 4;
  s.H.rowptr.resize(s.Natoms + 1);
  s.H.col.resize(nnzb);
  s.H.val.resize(16 * (size_t)nnzb);
  CHK(s, ltb_system_hamiltonian(s.sys, s.H.rowptr.data(), s.H.col.data(), s.H.val.data()));
  if (!s.opt.dump_dir.empty())  // MatView(H) in MATLAB format (:1504-1509)
    CHKIO(s, ltb_write_matlab_bsr(dump_path(s, "H.m").c_str(), "H", s.H.nrows, s.H.nrows, 4, s.H.rowptr.data(),
                                s.H.col.data(), s.H.val.data()));
  return LTB_OK;
}

PetscErrorCode initializeForSCF(State& s) {
  int nnzb, gnnz;
  double ne;
  CHK(s, ltb_system_sizes(s.sys, &s.Natoms, &nnzb, &gnnz, &ne));
  s.Gamma_ab.nrows = s.Natoms;
  s.Gamma_ab.bs = 1;
  s.Gamma_ab.rowptr.resize(s.Natoms + 1);
  s.Gamma_ab.col.resize(gnnz);
  s.Gamma_ab.val.resize(gnnz);
  CHK(s, ltb_system_gamma(s.sys, s.Gamma_ab.rowptr.data(), s.Gamma_ab.col.data(), s.Gamma_ab.val.data()));
  s.diagP.assign(s.nSize, 0.0);  // initializeDiagP (:1725-1740)
  if (!s.opt.dump_dir.empty())  // MatView(Gamma_ab) (:950-955)
    CHKIO(s, ltb_write_matlab_bsr(dump_path(s, "GAMMA_AB.m").c_str(), "GAMMA_AB", s.Natoms, s.Natoms, 1,
                                s.Gamma_ab.rowptr.data(), s.Gamma_ab.col.data(), s.Gamma_ab.val.data()));
  return LTB_OK;
}

PetscErrorCode GershGorinParallel(State& s, double* mn, double* mx) {
  CHK(s, ltb_gershgorin(s.ctx, mn, mx));
  return LTB_OK;
}

PetscErrorCode runKPM(State& s) {
  PhaseTimer timer(s.t_kpm, s.t_kpm_last, s.n_kpm);
  // bounds and rescale as :1541-1560
  double mn, mx;
  CHK(s, GershGorinParallel(s, &mn, &mx));
  s.a = (mx - mn) / 2;
  s.b = (mx + mn) / 2;
  CHK(s, ltb_rescale(s.ctx, s.a, s.b));
  if (s.opt.nrandoms > 0) {
    // stochastic moments of the trace, DOS and mu from its area (:1562-1701).
    // The reference then narrows [a, b] to where the DOS exceeds 0.001 per
    // orbital and rescales H with it (:1702-1712).  That can cut the band
    // edges, and T_m of an eigenvalue outside [-1, 1] grows like
    // cosh(m acosh|x|), so the Gershgorin frame is kept -- except in
    // reference-literal mode, which narrows as the reference does.
    const int nm = s.opt.dos_moments;
    std::vector<double> tr(nm);
    CHK(s, ltb_kpe_stochastic(s.ctx, nm - 1, s.opt.nrandoms, s.opt.seed + 1000003ULL * ++s.draws, s.opt.precision,
                              tr.data()));
    s.dosEnergies.assign(2 * nm, 0.0);
    s.dosValues.assign(2 * nm, 0.0);
    double mu, lo, hi;
    CHK(s, ltb_kpm_dos(tr.data(), nm, 2 * nm, s.a, s.b, s.nSize, s.Nelectrons, s.dosEnergies.data(),
                       s.dosValues.data(), &mu, &lo, &hi));
    if (!s.opt.dump_dir.empty())
      CHKIO(s, ltb_write_dos(dump_path(s, "dosfile.txt").c_str(), 2 * nm, s.dosEnergies.data(),
                             s.dosValues.data()));
    if (s.opt.reference_literal && hi > lo) {
      s.a = (hi - lo) / 2;
      s.b = (hi + lo) / 2;
      CHK(s, ltb_rescale(s.ctx, s.a, s.b));
    }
    s.mu_ha = mu;
    s.muval = (mu - s.b) / s.a;
    s.eband = 0;
    return LTB_OK;
  }
  // exact trace: mu by FP64 bisection on the exact moments in place of the
  // stochastic DOS search
  CHK(s, ltb_kpe_moments(s.ctx, s.opt.moments, s.opt.precision, nullptr, nullptr));
  const double kT = s.opt.kT_ev > 0 ? s.opt.kT_ev / 27.211386 : 0.0;
  ltb_fermi_params fp{s.opt.moments, kT, s.opt.quad_nodes, s.Nelectrons};
  ltb_density_result r;
  CHK(s, ltb_density(s.ctx, &fp, nullptr, &r));
  s.muval = r.mu;
  s.mu_ha = r.mu_ha;
  s.eband = r.eband;
  return LTB_OK;
}

PetscErrorCode getKPE(State& s, const Vec& in, Vec& out) {
  if ((int)in.size() != s.nSize) return LTB_ERR_ARG;
  std::vector<double> w(s.opt.moments);
  CHK(s, kpe_weights(s, w));
  out.assign(s.nSize, 0.0);
  CHK(s, ltb_kpe_apply(s.ctx, s.opt.moments, w.data(), 1, in.data(), out.data(), LTB_FP64));
  if (s.opt.kpe_scale != 1.0)
    for (double& x : out) x *= s.opt.kpe_scale;  // withSP2.cpp's VecScale(outVector, 2.0) (SP2:939)
  return LTB_OK;
}

PetscErrorCode getSP2(State& s, const Vec& in, Vec& out, int localsp2max, int currentsp2) {
  if (currentsp2 < 0) {
    std::fprintf(stderr, "some error with calling sp2 :: negative depth\n");  // :1208-1210
    return LTB_ERR_ARG;
  }
  if (currentsp2 == 0) return getKPE(s, in, out);
  Vec y, z;
  CHK(s, getSP2(s, in, y, localsp2max, currentsp2 - 1));
  CHK(s, getSP2(s, y, z, localsp2max, currentsp2 - 1));
  const double al = s.alphavals.at(currentsp2 - 1);
  out.resize(y.size());
  for (size_t i = 0; i < y.size(); i++) out[i] = (1.0 + al) * y[i] - al * z[i];
  return LTB_OK;
}

PetscErrorCode computeAlphas(State& s) {
  if (!s.opt.force_alphas.empty()) {
    const size_t n = s.opt.force_alphas.size();
    for (size_t k = 0; k < s.alphavals.size(); k++) s.alphavals[k] = s.opt.force_alphas[std::min(k, n - 1)];
    return LTB_OK;
  }
  if (s.opt.force_alpha != 0.0) {
    for (double& a : s.alphavals) a = s.opt.force_alpha;
    return LTB_OK;
  }
  for (int k = 0; k < s.opt.maxsp2; k++) {
    double tr = 0;
    if (s.opt.nrandoms > 0) {
      // the random-vector trace estimate of the level (:1240-1275)
      CHK(s, random_trace(s, k, &tr));
    } else if (k == 0) {
      // Tr P_KPE from the exact moments (the exact unit-vector trace)
      std::vector<double> d(s.nSize);
      CHK(s, ltb_density_diag(s.ctx, d.data()));
      for (double x : d) tr += s.opt.kpe_scale * x;
    } else {
      std::vector<double> d(s.nSize);
      CHK(s, diag_block(s, k, 0, s.nSize, d));
      for (double x : d) tr += x;
    }
    if (s.opt.verbose) std::printf("sp2count is %d, and trace is %lf \n", k, tr);  // :1277
    if (std::fabs(tr * 2 - s.Nelectrons) < 0.001) break;                           // :1278
    s.alphavals[k] = tr * 2 > s.Nelectrons ? -1.0 : 1.0;                           // :1280-1281
  }
  return LTB_OK;
}

PetscErrorCode computeDiagonalElements(State& s) {
  const int depth = s.opt.maxsp2 - 1;
  s.diagP.assign(s.nSize, 0.0);
  if (s.opt.nrandoms > 0)
    CHK(s, random_diag(s, depth, s.diagP));
  else if (depth == 0) {
    CHK(s, ltb_density_diag(s.ctx, s.diagP.data()));
    for (double& x : s.diagP) x *= s.opt.kpe_scale;
  } else
    CHK(s, diag_block(s, depth, 0, s.nSize, s.diagP));
  // chargeOnAtom = orbital2Atom . diagP (:1385)
  s.chargeOnAtom.assign(s.Natoms, 0.0);
  for (int i = 0; i < s.nSize; i++) s.chargeOnAtom[i / 4] += s.diagP[i];
  return LTB_OK;
}

PetscErrorCode computeDensityMatrix(State& s) {
  PhaseTimer timer(s.t_density, s.t_density_last, s.n_density);
  s.alphavals.assign(std::max(1, s.opt.maxsp2), 0.0);
  CHK(s, computeAlphas(s));
  CHK(s, computeDiagonalElements(s));
  return LTB_OK;
}

PetscErrorCode constructdeltaH(State& s) {
  // netCharges = populations - neutral; gammaq = Gamma . netCharges; the
  // shift enters H_total's diagonal on the device (:1424-1459)
  s.netCharges.resize(s.Natoms);
  for (int i = 0; i < s.Natoms; i++) s.netCharges[i] = s.mixedCharges[i] - s.neutralCharges[i];
  CHK(s, ltb_set_populations(s.ctx, s.mixedCharges.data()));
  s.gammaq.assign(s.Natoms, 0.0);
  CHK(s, ltb_get_shift(s.ctx, s.gammaq.data()));
  if (s.opt.reference_literal) {
    // the reference forms Delta H from the latest populations (:1418-1420) and
    // mixes Hamiltonians afterwards; report that Delta q and Gamma Delta q
    for (int i = 0; i < s.Natoms; i++) s.netCharges[i] = s.chargeOnAtom[i] - s.neutralCharges[i];
    for (int i = 0; i < s.Natoms; i++) {
      double g = 0;
      for (int k = s.Gamma_ab.rowptr[i]; k < s.Gamma_ab.rowptr[i + 1]; k++)
        g += s.Gamma_ab.val[k] * s.netCharges[s.Gamma_ab.col[k]];
      s.gammaq[i] = g;
    }
  }
  if (!s.opt.dump_dir.empty()) {
    // deltaH as 4x4 diagonal blocks, netCharges and gammaq (:1465-1484)
    std::vector<int> rp(s.Natoms + 1), col(s.Natoms);
    std::vector<double> blk(16 * (size_t)s.Natoms, 0.0);
    for (int i = 0; i < s.Natoms; i++) {
      rp[i + 1] = i + 1;
      col[i] = i;
      for (int a = 0; a < 4; a++) blk[16 * (size_t)i + 5 * a] = s.gammaq[i];
    }
    CHKIO(s, ltb_write_matlab_bsr(dump_path(s, "deltaH.m").c_str(), "deltaH", s.Natoms, s.Natoms, 4, rp.data(),
                                col.data(), blk.data()));
    CHKIO(s, ltb_write_matlab_vec(dump_path(s, "netCharge.m").c_str(), "Vec_netCharge", s.Natoms,
                                s.netCharges.data()));
    CHKIO(s, ltb_write_matlab_vec(dump_path(s, "GammaQ_vector.m").c_str(), "Vec_GammaQ", s.Natoms,
                                s.gammaq.data()));
  }
  return LTB_OK;
}

PetscErrorCode updateHtotal(State& s) {
  // linear mixing of the populations: mixed += mix (chargeOnAtom - mixed)
  for (int i = 0; i < s.Natoms; i++)
    s.mixedCharges[i] += s.opt.mix * (s.chargeOnAtom[i] - s.mixedCharges[i]);
  return LTB_OK;
}

// The reference's loop (main :1805-1840): one density of H before the loop,
// then per iteration Delta H, Hamiltonian mixing, runKPM, density, and the
// change |q_k - q_(k-1)| of successive outputs.
PetscErrorCode scf_literal(State& s, int* iterations, double* delta) {
  CHK(s, constructdeltaH(s));
  CHK(s, runKPM(s));
  CHK(s, computeDensityMatrix(s));
  print_phase_times(s);
  s.previousCharges = s.chargeOnAtom;
  int it = 0;
  double d = INFINITY;
  for (it = 0; it < s.opt.max_scf; it++) {
    CHK(s, updateHtotal(s));
    CHK(s, constructdeltaH(s));
    CHK(s, runKPM(s));
    CHK(s, computeDensityMatrix(s));
    d = 0;
    for (int i = 0; i < s.Natoms; i++) d = std::max(d, 2.0 * std::fabs(s.chargeOnAtom[i] - s.previousCharges[i]));
    if (s.opt.verbose) std::printf("scfiter : %d , the value of the scfDelta is %f \n", it, d);  // :1834
    s.previousCharges = s.chargeOnAtom;
    if (d < s.opt.tol) {
      it++;
      break;
    }
  }
  if (iterations) *iterations = it;
  if (delta) *delta = d;
  return d < s.opt.tol ? LTB_OK : LTB_NOT_CONVERGED;
}

// the same loop as scf on the device (ltb_scc_run) with the two-level SCC:
// most iterations at order two_level_order, corrected by the defect of the
// last full-order one; converged only on a full-order iteration
PetscErrorCode scf_two_level(State& s, int* iterations, double* delta) {
  if (s.opt.nrandoms > 0 || s.opt.maxsp2 > 1 || s.opt.two_level_order >= s.opt.moments || s.opt.kT_ev <= 0) {
    s.last_error = "the two-level SCC needs exact trace, maxsp2 1, Fermi-Dirac kT > 0 and an order below M";
    return LTB_ERR_ARG;
  }
  ltb_scc_params p{};
  p.M = s.opt.moments;
  p.kT = s.opt.kT_ev / 27.211386;
  p.K = s.opt.quad_nodes;
  p.mix = s.opt.mix;
  p.tol = s.opt.tol;
  p.max_iter = s.opt.max_scf;
  p.precision = s.opt.precision;
  p.switch_tol = s.opt.precision == LTB_FP32_ACC64 ? 1e-2 : 0.0;
  p.use_graph = 1;
  p.M_lo = s.opt.two_level_order;
  std::vector<double> pop(s.Natoms), hist(s.opt.max_scf);
  ltb_scc_result r{};
  const int st = ltb_scc_run(s.ctx, &p, s.mixedCharges.data(), pop.data(), hist.data(), &r);
  if (st && st != LTB_NOT_CONVERGED) {
    s.last_error = ltb_last_error(s.ctx);
    return st;
  }
  if (s.opt.verbose)
    for (int it = 0; it < r.iterations; it++)
      std::printf("scfiter : %d , the value of the scfDelta is %e \n", it, hist[it]);  // :1835
  s.chargeOnAtom = pop;
  CHK(s, ltb_get_mixed_populations(s.ctx, s.mixedCharges.data()));
  s.previousCharges = s.chargeOnAtom;
  s.mu_ha = r.mu_ha;
  s.eband = r.eband;
  if (iterations) *iterations = r.iterations;
  if (delta) *delta = r.dq;
  return st;
}

PetscErrorCode scf(State& s, int* iterations, double* delta) {
  if (s.opt.reference_literal) return scf_literal(s, iterations, delta);
  if (s.opt.two_level_order > 0) return scf_two_level(s, iterations, delta);
  int it = 0;
  double d = INFINITY;
  for (it = 0; it < s.opt.max_scf; it++) {
    CHK(s, constructdeltaH(s));
    CHK(s, runKPM(s));
    CHK(s, computeDensityMatrix(s));
    if (it == 0) print_phase_times(s);
    d = 0;
    for (int i = 0; i < s.Natoms; i++) d = std::max(d, 2.0 * std::fabs(s.chargeOnAtom[i] - s.mixedCharges[i]));
    if (s.opt.verbose) std::printf("scfiter : %d , the value of the scfDelta is %e \n", it, d);  // :1835
    s.previousCharges = s.chargeOnAtom;
    if (d < s.opt.tol) {
      it++;
      break;
    }
    CHK(s, updateHtotal(s));
  }
  if (iterations) *iterations = it;
  if (delta) *delta = d;
  return d < s.opt.tol ? LTB_OK : LTB_NOT_CONVERGED;
}

void clearAll(State& s) {
  if (s.ctx) ltb_destroy(s.ctx);
  if (s.sys) ltb_system_destroy(s.sys);
  s.ctx = nullptr;
  s.sys = nullptr;
}

}  // namespace ltb_petsc
