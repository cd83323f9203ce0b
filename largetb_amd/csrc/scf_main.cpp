// ltb_scf: the reference's driver (`./scf L M N nRandoms`, withSP2_deltaH.cpp
// main :1742-1849) over the entry points of ltb_petsc.hpp.  Without nRandoms
// the trace is exact over unit vectors; with it, runKPM's stochastic DOS and
// mu and the random-vector diagonal are used, as in the reference.
//   ltb_scf L M N [nRandoms] [--sigma A] [--seed S] [--moments M] [--kT eV] [--reference]
//                 [--maxsp2 D] [--tol Q] [--max-scf K] [--fp32] [--sp2-alpha a] [--dump DIR]
//                 [--mix m] [--reference-literal] [--sp2-alphas a1,a2,..] [--kpe-times-two]
//                 [--two-level M_lo]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/ltb_petsc.hpp"

using namespace ltb_petsc;

static double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s L M N [nRandoms] [--sigma A] [--seed S] [--moments M] [--kT eV] "
                 "[--reference] [--maxsp2 D] [--tol Q] [--max-scf K] [--fp32] [--sp2-alpha a] [--sp2-alphas a1,a2,..] [--kpe-times-two] [--dump DIR] [--mix m] "
                 "[--reference-literal] [--two-level M_lo]\n", argv[0]);
    return 2;
  }
  State s;
  s.opt.L = std::atoi(argv[1]);
  s.opt.M = std::atoi(argv[2]);
  s.opt.N = std::atoi(argv[3]);
  int first_opt = 4;
  if (argc > 4 && argv[4][0] != '-') {  // ./scf L M N nRandoms (:1747-1757)
    s.opt.nrandoms = std::atoi(argv[4]);
    first_opt = 5;
  }
  for (int i = first_opt; i < argc; i++) {
    std::string a = argv[i];
    auto val = [&]() { return i + 1 < argc ? argv[++i] : (char*)"0"; };
    if (a == "--sigma") s.opt.sigma_ang = std::atof(val());
    else if (a == "--seed") s.opt.seed = std::strtoull(val(), nullptr, 10);
    else if (a == "--moments") s.opt.moments = std::atoi(val());
    else if (a == "--kT") s.opt.kT_ev = std::atof(val());
    else if (a == "--reference") s.opt.lattice = LTB_LATTICE_REFERENCE;
    else if (a == "--maxsp2") s.opt.maxsp2 = std::atoi(val());
    else if (a == "--tol") s.opt.tol = std::atof(val());
    else if (a == "--max-scf") s.opt.max_scf = std::atoi(val());
    else if (a == "--fp32") s.opt.precision = LTB_FP32_ACC64;
    else if (a == "--sp2-alpha") s.opt.force_alpha = std::atof(val());
    else if (a == "--sp2-alphas") {  // comma-separated, one per SP2 level
      std::string list = val();
      for (size_t p = 0; p <= list.size();) {
        const size_t q = std::min(list.find(',', p), list.size());
        s.opt.force_alphas.push_back(std::atof(list.substr(p, q - p).c_str()));
        p = q + 1;
      }
    } else if (a == "--kpe-times-two") s.opt.kpe_scale = 2.0;
    else if (a == "--dump") s.opt.dump_dir = val();
    else if (a == "--mix") s.opt.mix = std::atof(val());
    else if (a == "--reference-literal") s.opt.reference_literal = true;
    else if (a == "--two-level") s.opt.two_level_order = std::atoi(val());
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  double t0 = now();
  std::printf("starting to write coordinates\n M  L  N %d %d %d\n", s.opt.M, s.opt.L, s.opt.N);
  int st = arrangeCoordinates(s);
  if (st) return std::fprintf(stderr, "arrangeCoordinates: %s\n", s.last_error.c_str()), st;
  std::printf("time arranging Coordinates : %g\n", now() - t0);
  t0 = now();
  if ((st = initializeForH(s)) || (st = constructH(s)) || (st = initializeForSCF(s)))
    return std::fprintf(stderr, "setup: %s\n", s.last_error.c_str()), st;
  std::printf("time Hamiltonian : %g\n", now() - t0);
  t0 = now();
  int iters = 0;
  double delta = 0;
  st = scf(s, &iters, &delta);
  const double t_scf = now() - t0;
  if (st && st != LTB_NOT_CONVERGED) return std::fprintf(stderr, "scf: %s\n", s.last_error.c_str()), st;
  double qsi_min = 1e300, qsi_max = -1e300, qo_min = 1e300, qo_max = -1e300, tot = 0;
  std::vector<int> Z(s.Natoms);
  ltb_system_coords(s.sys, Z.data(), nullptr);
  for (int i = 0; i < s.Natoms; i++) {
    double q = 2.0 * s.chargeOnAtom[i];
    tot += q;
    if (Z[i] == 14) qsi_min = std::min(qsi_min, q), qsi_max = std::max(qsi_max, q);
    else qo_min = std::min(qo_min, q), qo_max = std::max(qo_max, q);
  }
  std::printf("{\"status\": %d, \"atoms\": %d, \"iterations\": %d, \"delta\": %.3e, \"seconds\": %.3f, "
              "\"mu_ha\": %.12f, \"eband_ha\": %.12f, \"total_charge\": %.12f, "
              "\"q_si\": [%.9f, %.9f], \"q_o\": [%.9f, %.9f], "
              "\"runKPM\": {\"calls\": %d, \"seconds\": %.4f}, "
              "\"computeDensityMatrix\": {\"calls\": %d, \"seconds\": %.4f}}\n",
              st, s.Natoms, iters, delta, t_scf, s.mu_ha, s.eband, tot, qsi_min, qsi_max,
              qo_min, qo_max, s.n_kpm, s.t_kpm, s.n_density, s.t_density);
  // full charge vector for checkers
  if (const char* path = std::getenv("LTB_SCF_CHARGES")) {
    if (FILE* f = std::fopen(path, "w")) {
      for (int i = 0; i < s.Natoms; i++) std::fprintf(f, "%.17g\n", 2.0 * s.chargeOnAtom[i]);
      std::fclose(f);
    }
  }
  clearAll(s);
  return st;
}
