// Host-side assembly of one alpha-quartz tight-binding system:
//   geometry        (initializeCoordinates, withSP2_deltaH.cpp:504-667)
//   neighbour list  (getHaloAtoms :372-501, ModifyNList :355-370) -- here an
//                   O(N) cell list with minimum images for large periodic
//                   cells, and the reference's image-by-image scan otherwise
//   H0 blocks       (get4by4HmatrixElements :211-349, constructH :969-1066)
//   Gamma           (initializeForSCF :854-900)
#include "system.hpp"

#include "quartz_model.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <set>

namespace ltb {
namespace {

const double kFrac[9][3] = LTB_QUARTZ_FRAC;
const int kZCell[9] = LTB_QUARTZ_Z;
const double kSK[3][20][4] = LTB_SK_TABLES;  // [0] Si-Si, [1] O-O, [2] Si-O

// SplitMix64 stream -> Box-Muller pairs: a portable, documented generator so
// the displaced structures are the same on every platform and compiler.
struct Gauss {
  uint64_t s;
  bool have = false;
  double spare = 0;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)((next() >> 11) + 1) * 0x1.0p-53; }  // (0, 1]
  double normal() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = uniform(), u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1)), t = 2.0 * kPi * u2;
    spare = r * std::sin(t);
    have = true;
    return r * std::cos(t);
  }
};

struct Neighbor {
  int j;
  double d[3];  // r_j - r_i (image included)
};

Status fail(int code, std::string msg) { return Status{code, std::move(msg)}; }

// inverse of a 3x3 matrix whose rows are the cell vectors
bool invert3(const double a[9], double inv[9]) {
  double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8],
         c02 = a[3] * a[7] - a[4] * a[6];
  double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(std::fabs(det) > 1e-12)) return false;
  double id = 1.0 / det;
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

// perpendicular width of the cell along row d
double width(const double c[9], int d) {
  int e = (d + 1) % 3, f = (d + 2) % 3;
  double cr[3] = {c[3 * e + 1] * c[3 * f + 2] - c[3 * e + 2] * c[3 * f + 1],
                  c[3 * e + 2] * c[3 * f + 0] - c[3 * e + 0] * c[3 * f + 2],
                  c[3 * e + 0] * c[3 * f + 1] - c[3 * e + 1] * c[3 * f + 0]};
  double vol = std::fabs(c[3 * d] * cr[0] + c[3 * d + 1] * cr[1] + c[3 * d + 2] * cr[2]);
  return vol / std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
}

constexpr double kRcut2 = 16.0;  // r^2 < 16 Bohr^2 (:392)
constexpr double kRmin2 = 0.1;   // r^2 > 0.1 (:421)

// Reference scan: every image (lp, mp, np) in loop order, every atom j.
// Repeated j keeps the last image (INSERT_VALUES, constructH :1022).
Status scan_images(const ltb_system& s, std::vector<std::vector<Neighbor>>& nb) {
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) lo[d] = hi[d] = 0;
  for (int d = 0; d < 3; d++)
    if (s.periodic_mask >> d & 1) lo[d] = -1, hi[d] = 1;
  const int n = s.natoms;
  for (int i = 0; i < n; i++) {
    std::vector<Neighbor>& out = nb[i];
    int entries = 0;
    for (int l = lo[0]; l <= hi[0]; l++)
      for (int m = lo[1]; m <= hi[1]; m++)
        for (int q = lo[2]; q <= hi[2]; q++) {
          double sh[3];
          for (int d = 0; d < 3; d++) sh[d] = l * s.cell[d] + m * s.cell[3 + d] + q * s.cell[6 + d];
          for (int j = 0; j < n; j++) {
            double dx = s.xyz[3 * j] + sh[0] - s.xyz[3 * i];
            double dy = s.xyz[3 * j + 1] + sh[1] - s.xyz[3 * i + 1];
            double dz = s.xyz[3 * j + 2] + sh[2] - s.xyz[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 <= kRmin2 && !(j == i && l == 0 && m == 0 && q == 0))
              return fail(2, "atoms " + std::to_string(i) + " and " + std::to_string(j) +
                                 " closer than sqrt(0.1) Bohr");
            if (r2 < kRcut2 && r2 > kRmin2) {
              if (++entries > kMaxNeighbors)
                return fail(3, "atom " + std::to_string(i) + " has more than 20 neighbours");
              auto it = std::find_if(out.begin(), out.end(), [&](const Neighbor& x) { return x.j == j; });
              Neighbor nbv{j, {dx, dy, dz}};
              if (it != out.end())
                *it = nbv;
              else
                out.push_back(nbv);
            }
          }
        }
  }
  return Status::ok();
}

// Cell list with minimum images; valid when every periodic direction is
// wider than twice the cutoff (then at most one image of j is in range).
Status cell_list(const ltb_system& s, const double inv[9], std::vector<std::vector<Neighbor>>& nb) {
  const int n = s.natoms;
  std::vector<double> f(3 * (size_t)n);
  double fmin[3] = {1e300, 1e300, 1e300}, fmax[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      double v = s.xyz[3 * i] * inv[d] + s.xyz[3 * i + 1] * inv[3 + d] + s.xyz[3 * i + 2] * inv[6 + d];
      f[3 * i + d] = v;
      fmin[d] = std::min(fmin[d], v);
      fmax[d] = std::max(fmax[d], v);
    }
  int nbin[3];
  double w[3], f0[3], span[3];
  for (int d = 0; d < 3; d++) {
    w[d] = width(s.cell, d);
    bool per = s.periodic_mask >> d & 1;
    f0[d] = per ? 0.0 : fmin[d];
    span[d] = per ? 1.0 : std::max(fmax[d] - fmin[d], 1e-9);
    nbin[d] = std::max(1, (int)std::floor(w[d] * span[d] / 4.0));
    nbin[d] = std::min(nbin[d], 1024);
  }
  auto bin_of = [&](int i, int d) {
    bool per = s.periodic_mask >> d & 1;
    double v = (f[3 * i + d] - f0[d]) / span[d];
    if (per) v -= std::floor(v);
    int b = (int)std::floor(v * nbin[d]);
    return std::min(std::max(b, 0), nbin[d] - 1);
  };
  size_t nb_tot = (size_t)nbin[0] * nbin[1] * nbin[2];
  std::vector<int> head(nb_tot + 1, 0), items(n);
  std::vector<int> bid(n);
  for (int i = 0; i < n; i++) {
    bid[i] = (bin_of(i, 0) * nbin[1] + bin_of(i, 1)) * nbin[2] + bin_of(i, 2);
    head[bid[i] + 1]++;
  }
  for (size_t b = 0; b < nb_tot; b++) head[b + 1] += head[b];
  {
    std::vector<int> fill(head.begin(), head.end() - 1);
    for (int i = 0; i < n; i++) items[fill[bid[i]]++] = i;
  }
  for (int i = 0; i < n; i++) {
    int bi[3] = {bin_of(i, 0), bin_of(i, 1), bin_of(i, 2)};
    std::set<size_t> bins;
    for (int o0 = -1; o0 <= 1; o0++)
      for (int o1 = -1; o1 <= 1; o1++)
        for (int o2 = -1; o2 <= 1; o2++) {
          int b[3] = {bi[0] + o0, bi[1] + o1, bi[2] + o2};
          bool ok = true;
          for (int d = 0; d < 3; d++) {
            if (s.periodic_mask >> d & 1)
              b[d] = ((b[d] % nbin[d]) + nbin[d]) % nbin[d];
            else if (b[d] < 0 || b[d] >= nbin[d])
              ok = false;
          }
          if (ok) bins.insert(((size_t)b[0] * nbin[1] + b[1]) * nbin[2] + b[2]);
        }
    std::vector<Neighbor>& out = nb[i];
    for (size_t b : bins)
      for (int k = head[b]; k < head[b + 1]; k++) {
        int j = items[k];
        if (j == i) continue;
        double df[3];
        for (int d = 0; d < 3; d++) {
          df[d] = f[3 * j + d] - f[3 * i + d];
          if (s.periodic_mask >> d & 1) df[d] -= std::nearbyint(df[d]);
        }
        double dv[3];
        for (int d = 0; d < 3; d++) dv[d] = df[0] * s.cell[d] + df[1] * s.cell[3 + d] + df[2] * s.cell[6 + d];
        double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
        if (r2 <= kRmin2)
          return fail(2, "atoms " + std::to_string(i) + " and " + std::to_string(j) +
                             " closer than sqrt(0.1) Bohr");
        if (r2 < kRcut2) {
          if ((int)out.size() >= kMaxNeighbors)
            return fail(3, "atom " + std::to_string(i) + " has more than 20 neighbours");
          out.push_back(Neighbor{j, {dv[0], dv[1], dv[2]}});
        }
      }
  }
  return Status::ok();
}

}  // namespace

// get4by4HmatrixElements (:211-349).  The radial sums
// S_c(r) = sum_{p=1..20} tab[p][c] * cos(p acos(r/10)) * exp(-p r)
// are formed once per pair; the angular factors follow :224-253.
Status sk_block(int zi, int zj, double x, double y, double z, double h[16]) {
  double r2 = x * x + y * y + z * z, r = std::sqrt(r2);
  const double(*tab)[4] = nullptr;
  for (int k = 0; k < 16; k++) h[k] = 0;
  if ((zi == 14 && zj == 8) || (zi == 8 && zj == 14)) {
    if (r < 0.01) return fail(2, "Si and O closer than 0.01 Bohr");
    tab = kSK[2];
  } else if (zi == 14 && zj == 14) {
    if (r < 0.01) {  // onsite (:258-263)
      h[0] = LTB_ONSITE_SI_S;
      h[5] = h[10] = h[15] = LTB_ONSITE_SI_P;
      return Status::ok();
    }
    tab = kSK[0];
  } else if (zi == 8 && zj == 8) {
    if (r < 0.01) {  // onsite (:299-304)
      h[0] = LTB_ONSITE_O_S;
      h[5] = h[10] = h[15] = LTB_ONSITE_O_P;
      return Status::ok();
    }
    tab = kSK[1];
  } else {
    return fail(2, "unknown atom pair " + std::to_string(zi) + "-" + std::to_string(zj));
  }
  double S[4] = {0, 0, 0, 0};
  const double u = std::acos(r / 10);
  for (int p = 1; p <= 20; p++) {
    double t = std::cos(p * u) * std::exp(-p * r);
    for (int c = 0; c < 4; c++) S[c] += tab[p - 1][c] * t;
  }
  const double d[3] = {x, y, z};
  h[0] = S[3];
  for (int a = 0; a < 3; a++) {
    h[1 + a] = S[2] * d[a] / r;
    h[4 * (1 + a)] = -S[2] * d[a] / r;
    for (int b = 0; b < 3; b++) {
      double dd = d[a] * d[b] / r2;
      h[4 * (1 + a) + 1 + b] = S[0] * dd - S[1] * (dd - (a == b ? 1.0 : 0.0));
    }
  }
  return Status::ok();
}

Status build_quartz(int L, int M, int N, int lattice, double sigma_ang, uint64_t seed,
                    ltb_system& s) {
  if (L <= 0 || M <= 0 || N <= 0 || (long long)L * M * N * 9 > (1LL << 28))
    return fail(1, "bad supercell size");
  if (lattice != 0 && lattice != 1) return fail(1, "bad lattice mode");
  double a, c;
  if (lattice == 1) {
    a = 9.725;  // Bohr (:508)
    c = 10.7;
  } else {
    a = 4.913 * kAngToBohr;
    c = 5.405 * kAngToBohr;
  }
  const double R[3][3] = {{a, 0.0, 0.0}, {-a * std::sin(kPi / 6), a * std::cos(kPi / 6), 0.0}, {0.0, 0.0, c}};
  s.natoms = 9 * L * M * N;
  s.Z.assign(s.natoms, 0);
  s.xyz.assign(3 * (size_t)s.natoms, 0.0);
  int at = 0;
  for (int i = 0; i < L; i++)
    for (int j = 0; j < M; j++)
      for (int k = 0; k < N; k++) {
        double o[3];
        for (int d = 0; d < 3; d++) o[d] = i * R[0][d] + j * R[1][d] + k * R[2][d];
        for (int t = 0; t < 9; t++, at++) {
          for (int d = 0; d < 3; d++)
            s.xyz[3 * at + d] = o[d] + (kFrac[t][0] * R[0][d] + kFrac[t][1] * R[1][d] + kFrac[t][2] * R[2][d]);
          s.Z[at] = kZCell[t];
        }
      }
  if (lattice == 0 && sigma_ang > 0) {
    Gauss g{seed};
    for (size_t q = 0; q < s.xyz.size(); q++) s.xyz[q] += sigma_ang * kAngToBohr * g.normal();
  }
  for (int d = 0; d < 3; d++) {
    s.cell[d] = L * R[0][d];
    s.cell[3 + d] = M * R[1][d];
    s.cell[6 + d] = N * R[2][d];
  }
  // reference lattice: no periodic images along z (:400-403)
  s.periodic_mask = lattice == 1 ? 3 : 7;
  return build_from_coords(s);
}

Status build_from_coords(ltb_system& s) {
  const int n = s.natoms;
  if (n <= 0 || (int)s.Z.size() != n || s.xyz.size() != 3 * (size_t)n) return fail(1, "bad coordinates");
  for (int i = 0; i < n; i++)
    if (s.Z[i] != 14 && s.Z[i] != 8) return fail(2, "atomic number must be 14 (Si) or 8 (O)");
  for (size_t q = 0; q < s.xyz.size(); q++)
    if (!std::isfinite(s.xyz[q])) return fail(4, "non-finite coordinate");
  double inv[9];
  bool have_inv = invert3(s.cell, inv);
  if (s.periodic_mask && !have_inv) return fail(1, "singular periodic cell");
  bool large = have_inv;
  for (int d = 0; d < 3; d++)
    if ((s.periodic_mask >> d & 1) && width(s.cell, d) <= 2.0 * 4.0 + 1e-9) large = false;
  std::vector<std::vector<Neighbor>> nb(n);
  Status st = large ? cell_list(s, inv, nb) : scan_images(s, nb);
  if (st.code) return st;

  s.h_rowptr.assign(n + 1, 0);
  s.g_rowptr.assign(n + 1, 0);
  s.h_col.clear();
  s.h_val.clear();
  s.g_col.clear();
  s.g_val.clear();
  double blk[16];
  for (int i = 0; i < n; i++) {
    std::vector<Neighbor>& v = nb[i];
    std::sort(v.begin(), v.end(), [](const Neighbor& x, const Neighbor& y) { return x.j < y.j; });
    st = sk_block(s.Z[i], s.Z[i], 0, 0, 0, blk);
    if (st.code) return st;
    s.h_col.push_back(i);
    s.h_val.insert(s.h_val.end(), blk, blk + 16);
    s.g_col.push_back(i);
    s.g_val.push_back(-1.0);  // -Us, Us = 1.0 (:876, :887)
    for (const Neighbor& q : v) {
      st = sk_block(s.Z[i], s.Z[q.j], q.d[0], q.d[1], q.d[2], blk);
      if (st.code) return st;
      s.h_col.push_back(q.j);
      s.h_val.insert(s.h_val.end(), blk, blk + 16);
      double r = std::sqrt(q.d[0] * q.d[0] + q.d[1] * q.d[1] + q.d[2] * q.d[2]);
      s.g_col.push_back(q.j);
      s.g_val.push_back(1.0 / r);  // (:890-896)
    }
    s.h_rowptr[i + 1] = (int)s.h_col.size();
    s.g_rowptr[i + 1] = (int)s.g_col.size();
  }
  s.p0.resize(n);
  s.nelectrons = 0;
  for (int i = 0; i < n; i++) {
    s.p0[i] = s.Z[i] == 14 ? 2.0 : 3.0;  // per-spin valence: Si 4/2, O 6/2
    s.nelectrons += 2.0 * s.p0[i];
  }
  return Status::ok();
}
Status pairs_within(const ltb_system& s, double rc, std::vector<int>& rowptr,
                    std::vector<int>& nbj, std::vector<double>& nbd) {
  const int n = s.natoms;
  double inv[9];
  const bool have_inv = invert3(s.cell, inv);
  if (s.periodic_mask && !have_inv) return fail(1, "singular periodic cell");
  const double rc2 = rc * rc;
  rowptr.assign(n + 1, 0);
  nbj.clear();
  nbd.clear();
  int reach[3] = {0, 0, 0};
  bool minimum_image = have_inv;
  for (int d = 0; d < 3; d++)
    if (s.periodic_mask >> d & 1) {
      double w = width(s.cell, d);
      reach[d] = (int)std::ceil(rc / w) + 1;  // positions may sit anywhere in the cell
      if (w <= 2.0 * rc) minimum_image = false;
    }
  if (minimum_image && n > 64) {
    // binned search, one image per pair (every periodic width exceeds 2 rc)
    std::vector<double> f(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++)
        f[3 * i + d] = s.xyz[3 * i] * inv[d] + s.xyz[3 * i + 1] * inv[3 + d] + s.xyz[3 * i + 2] * inv[6 + d];
    int nbin[3];
    double fmin[3], span[3];
    for (int d = 0; d < 3; d++) {
      bool per = s.periodic_mask >> d & 1;
      double lo = 1e300, hi = -1e300;
      for (int i = 0; i < n; i++) lo = std::min(lo, f[3 * i + d]), hi = std::max(hi, f[3 * i + d]);
      fmin[d] = per ? 0.0 : lo;
      span[d] = per ? 1.0 : std::max(hi - lo, 1e-9);
      nbin[d] = std::max(1, std::min(512, (int)std::floor(width(s.cell, d) * span[d] / rc)));
    }
    auto bin_of = [&](int i, int d) {
      double v = (f[3 * i + d] - fmin[d]) / span[d];
      if (s.periodic_mask >> d & 1) v -= std::floor(v);
      return std::min(std::max((int)std::floor(v * nbin[d]), 0), nbin[d] - 1);
    };
    size_t nb_tot = (size_t)nbin[0] * nbin[1] * nbin[2];
    std::vector<int> head(nb_tot + 1, 0), items(n), bid(n);
    for (int i = 0; i < n; i++) {
      bid[i] = (bin_of(i, 0) * nbin[1] + bin_of(i, 1)) * nbin[2] + bin_of(i, 2);
      head[bid[i] + 1]++;
    }
    for (size_t b = 0; b < nb_tot; b++) head[b + 1] += head[b];
    {
      std::vector<int> fill(head.begin(), head.end() - 1);
      for (int i = 0; i < n; i++) items[fill[bid[i]]++] = i;
    }
    for (int i = 0; i < n; i++) {
      int bi[3] = {bin_of(i, 0), bin_of(i, 1), bin_of(i, 2)};
      std::set<size_t> bins;
      for (int o0 = -1; o0 <= 1; o0++)
        for (int o1 = -1; o1 <= 1; o1++)
          for (int o2 = -1; o2 <= 1; o2++) {
            int b[3] = {bi[0] + o0, bi[1] + o1, bi[2] + o2};
            bool ok = true;
            for (int d = 0; d < 3; d++) {
              if (s.periodic_mask >> d & 1)
                b[d] = ((b[d] % nbin[d]) + nbin[d]) % nbin[d];
              else if (b[d] < 0 || b[d] >= nbin[d])
                ok = false;
            }
            if (ok) bins.insert(((size_t)b[0] * nbin[1] + b[1]) * nbin[2] + b[2]);
          }
      for (size_t b : bins)
        for (int k = head[b]; k < head[b + 1]; k++) {
          int j = items[k];
          if (j == i) continue;
          double df[3], dv[3];
          for (int d = 0; d < 3; d++) {
            df[d] = f[3 * j + d] - f[3 * i + d];
            if (s.periodic_mask >> d & 1) df[d] -= std::nearbyint(df[d]);
          }
          for (int d = 0; d < 3; d++) dv[d] = df[0] * s.cell[d] + df[1] * s.cell[3 + d] + df[2] * s.cell[6 + d];
          double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
          if (r2 < rc2) {
            nbj.push_back(j);
            nbd.insert(nbd.end(), dv, dv + 3);
          }
        }
      rowptr[i + 1] = (int)nbj.size();
    }
    return Status::ok();
  }
  // every image within reach, every atom (small cells)
  for (int i = 0; i < n; i++) {
    for (int l = -reach[0]; l <= reach[0]; l++)
      for (int m = -reach[1]; m <= reach[1]; m++)
        for (int q = -reach[2]; q <= reach[2]; q++) {
          double sh[3];
          for (int d = 0; d < 3; d++) sh[d] = l * s.cell[d] + m * s.cell[3 + d] + q * s.cell[6 + d];
          for (int j = 0; j < n; j++) {
            if (j == i && l == 0 && m == 0 && q == 0) continue;
            double dv[3] = {s.xyz[3 * j] + sh[0] - s.xyz[3 * i], s.xyz[3 * j + 1] + sh[1] - s.xyz[3 * i + 1],
                            s.xyz[3 * j + 2] + sh[2] - s.xyz[3 * i + 2]};
            double r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            if (r2 < rc2) {
              nbj.push_back(j);
              nbd.insert(nbd.end(), dv, dv + 3);
            }
          }
        }
    rowptr[i + 1] = (int)nbj.size();
  }
  return Status::ok();
}
// Visiting order of the atoms for the Chebyshev sweep.  Atoms are binned on a
// grid of ~9 Bohr bins along the cell rows (fractional coordinates; Cartesian
// bounding box when the cell is singular).  With p the axis of most bins, r
// the axis of fewest and q the third, the order runs over strips of w q-bins:
//   (q / w, p, q % w, r, atom index)
// so a bond along r is a few atoms apart in the order, along q ~ nb_r bins
// and along p ~ w * nb_r bins.  w is picked to keep that last distance near
// 450 atoms: the t_n rows re-read by neighbours then stay within one XCD's L2
// for a 256-column group (4 KB of FP32 rows per atom).
Status locality_order(const ltb_system& s, std::vector<int>& order) {
  const int n = s.natoms;
  order.resize(n);
  for (int i = 0; i < n; i++) order[i] = i;
  if (n < 64) return Status::ok();
  double inv[9];
  const bool frac = invert3(s.cell, inv);
  std::vector<double> f(3 * (size_t)n);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      const double* r = &s.xyz[3 * (size_t)i];
      double v = frac ? r[0] * inv[d] + r[1] * inv[3 + d] + r[2] * inv[6 + d] : r[d];
      if (frac && (s.periodic_mask >> d & 1)) v -= std::floor(v);
      f[3 * (size_t)i + d] = v;
      lo[d] = std::min(lo[d], v);
      hi[d] = std::max(hi[d], v);
    }
  int nb[3];
  for (int d = 0; d < 3; d++) {
    const bool per = frac && (s.periodic_mask >> d & 1);
    const double span = per ? 1.0 : std::max(hi[d] - lo[d], 1e-9);
    if (per) lo[d] = 0.0;
    const double extent = frac ? span * width(s.cell, d) : span;
    nb[d] = std::min(1024, std::max(1, (int)std::floor(extent / 9.0)));
    hi[d] = span;
  }
  int ax[3] = {0, 1, 2};
  std::sort(ax, ax + 3, [&](int a, int b) { return nb[a] > nb[b] || (nb[a] == nb[b] && a < b); });
  const int p = ax[0], q = ax[1], r = ax[2];
  const double per_bin = (double)n / ((double)nb[0] * nb[1] * nb[2]);
  const int w = std::max(1, std::min(nb[q], (int)std::lround(450.0 / (nb[r] * per_bin))));
  auto bin = [&](int i, int d) {
    int b = (int)std::floor((f[3 * (size_t)i + d] - lo[d]) / hi[d] * nb[d]);
    return std::min(nb[d] - 1, std::max(0, b));
  };
  std::vector<long long> key(n);
  for (int i = 0; i < n; i++) {
    const long long bq = bin(i, q), bp = bin(i, p), br = bin(i, r);
    key[i] = (((bq / w) * nb[p] + bp) * w + bq % w) * nb[r] + br;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  return Status::ok();
}
void gaussian_fill(uint64_t seed, size_t n, double* out) {
  Gauss g{seed};
  for (size_t k = 0; k < n; k++) out[k] = g.normal();
}

void cone_ranges(const std::vector<char>& in, int nat, int max_ranges, std::vector<std::pair<int, int>>& out) {
  out.clear();
  for (int a = 0; a < nat;) {
    if (!in[a]) {
      a++;
      continue;
    }
    int b = a;
    while (b < nat && in[b]) b++;
    out.push_back({a, b - a});
    a = b;
  }
  if (out.size() > 1 && out.front().first == 0 && out.back().first + out.back().second == nat) {
    out.back().second += out.front().second;  // a run across the end of the atom order
    out.erase(out.begin());
  }
  for (long long gap = 8; (int)out.size() > std::max(1, max_ranges); gap *= 2) {
    std::vector<std::pair<int, int>> m;
    for (const auto& r : out) {
      if (!m.empty()) {
        const int end = m.back().first + m.back().second;  // may pass nat only for the last run
        if (r.first - end <= gap) {
          m.back().second = r.first + r.second - m.back().first;
          continue;
        }
      }
      m.push_back(r);
    }
    out.swap(m);
  }
  long long tot = 0;
  for (const auto& r : out) tot += r.second;
  if (tot >= nat) out.assign(1, {0, nat});
}

void light_cones(int nat, const int* rowptr, const int* col, int apc, int max_ranges,
                 std::vector<std::vector<std::pair<int, int>>>& idx, std::vector<std::pair<int, int>>& ranges) {
  idx.clear();
  ranges.clear();
  std::vector<char> in(nat);
  std::vector<int> frontier, next;
  std::vector<std::pair<int, int>> rs;
  for (int a0 = 0; a0 < nat; a0 += apc) {
    const int a1 = std::min(a0 + apc, nat);
    std::fill(in.begin(), in.end(), 0);
    frontier.clear();
    for (int a = a0; a < a1; a++) {
      in[a] = 1;
      frontier.push_back(a);
    }
    std::vector<std::pair<int, int>> hops;
    for (;;) {
      cone_ranges(in, nat, max_ranges, rs);
      hops.push_back({(int)ranges.size(), (int)rs.size()});
      ranges.insert(ranges.end(), rs.begin(), rs.end());
      if (frontier.empty() || (rs.size() == 1 && rs[0].second >= nat)) break;
      next.clear();
      for (int a : frontier)
        for (int k = rowptr[a]; k < rowptr[a + 1]; k++) {
          const int b = col[k];
          if (!in[b]) {
            in[b] = 1;
            next.push_back(b);
          }
        }
      frontier.swap(next);
    }
    idx.push_back(std::move(hops));
  }
}

}  // namespace ltb
