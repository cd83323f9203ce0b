// Host-side structure of one tight-binding system: geometry, neighbour list,
// H0 as BSR 4x4 (FP64) and the Gamma matrix (CSR, FP64).
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

struct ltb_system {
  int natoms = 0;
  std::vector<int> Z;         // 14 Si / 8 O
  std::vector<double> xyz;    // Bohr, [natoms][3]
  double cell[9] = {0};       // periodic vectors (rows), Bohr
  int periodic_mask = 0;
  // H0: BSR, diagonal block first in each row, then ascending columns
  std::vector<int> h_rowptr, h_col;
  std::vector<double> h_val;  // [nnzb][16], row-major 4x4
  // Gamma: CSR, same row layout
  std::vector<int> g_rowptr, g_col;
  std::vector<double> g_val;
  double nelectrons = 0;
  std::vector<double> p0;     // neutral per-spin populations
};

namespace ltb {
constexpr int kMaxNeighbors = 20;  // maxN_A, withSP2_deltaH.cpp:26
constexpr double kAngToBohr = 1.88973;
constexpr double kPi = 3.141592653589793238463;

// status + message; status 0 = OK
struct Status {
  int code = 0;
  std::string msg;
  static Status ok() { return {}; }
};

Status build_quartz(int L, int M, int N, int lattice, double sigma_ang, uint64_t seed,
                    ltb_system& s);
Status build_from_coords(ltb_system& s);  // uses s.Z, s.xyz, s.cell, s.periodic_mask
Status sk_block(int zi, int zj, double x, double y, double z, double h[16]);

// All pairs (i, j, image) with 0 < |r_j + image - r_i| < rc, images along the
// periodic cell rows as far as rc reaches.  Per atom i: entries
// [rowptr[i], rowptr[i+1]) hold j and the displacement r_j + image - r_i.
Status pairs_within(const ltb_system& s, double rc, std::vector<int>& rowptr,
                    std::vector<int>& nbj, std::vector<double>& nbd);

// Visiting order of the atoms that keeps bonded atoms close in sequence
// (strips of spatial bins); a permutation of 0..natoms-1.
Status locality_order(const ltb_system& s, std::vector<int>& order);

// n standard normal numbers of the SplitMix64 -> Box-Muller stream of `seed`
// (the generator of the structure displacements)
void gaussian_fill(uint64_t seed, size_t n, double* out);

// The atoms marked in `in` (nat entries) as at most max_ranges cyclic ranges
// {lo, len}: its runs of consecutive atoms (a run across the end of the order
// wraps: lo + len > nat), merged across the smallest gaps until few enough;
// {0, nat} when they hold every atom.  Every marked atom lies in a range.
void cone_ranges(const std::vector<char>& in, int nat, int max_ranges, std::vector<std::pair<int, int>>& out);

// Light cones over the block graph of H (rowptr / col, nat block rows): for
// chunk k of the atoms (k*apc .. min((k+1)*apc, nat) - 1) and hop count
// h = 0, 1, ..., the atoms within h hops of the chunk's atoms, as
// cone_ranges gives them.  idx[k][h] = {first, count} into `ranges`; a chunk's
// list stops once its cone holds every atom or stops growing, and its last
// entry stands for every larger h.
void light_cones(int nat, const int* rowptr, const int* col, int apc, int max_ranges,
                 std::vector<std::vector<std::pair<int, int>>>& idx, std::vector<std::pair<int, int>>& ranges);
}  // namespace ltb
