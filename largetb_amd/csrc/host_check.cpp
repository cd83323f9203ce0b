// Host-only self-check of the structure assembly (system.cpp) and the MATLAB
// dump reader/writer (matlab_io.cpp), built with AddressSanitizer and
// UndefinedBehaviorSanitizer by largetb_amd/build.py (build_host_check) and
// run by the CPU tests.  It drives every host path: the quartz builder in
// both lattices, the cell list and the image-by-image scan, a cluster
// without a cell, the error paths (atoms on top of each other, too many
// neighbours), the descriptor pair search, the visiting order, and the MATLAB
// files (round trip, truncated and malformed input), and the light-cone
// tables of the Chebyshev sweep (against an independent distance search on
// quartz, ring and random graphs).  Prints "host_check ok" and exits 0 when
// every check holds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/ltb.h"
#include "system.hpp"

namespace {

int g_fail = 0;

void check(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "host_check: FAILED %s\n", what.c_str());
    g_fail++;
  }
}

// the neighbour pattern and blocks of two assemblies agree
bool same_system(const ltb_system& a, const ltb_system& b) {
  if (a.h_rowptr != b.h_rowptr || a.h_col != b.h_col || a.g_col != b.g_col) return false;
  for (size_t k = 0; k < a.h_val.size(); k++)
    if (std::fabs(a.h_val[k] - b.h_val[k]) > 1e-12) return false;
  return true;
}

void quartz_paths() {
  for (int lattice : {LTB_LATTICE_CONFIG, LTB_LATTICE_REFERENCE})
    for (int n : {1, 2, 3, 5}) {
      ltb_system s;
      ltb::Status st = ltb::build_quartz(n, n, n == 1 ? 2 : n, lattice, 0.08, 7, s);
      check(st.code == 0, "build_quartz " + std::to_string(n) + " lattice " + std::to_string(lattice) + ": " + st.msg);
      if (st.code) continue;
      // rows: diagonal block first, then strictly ascending columns
      for (int i = 0; i < s.natoms; i++) {
        check(s.h_col[s.h_rowptr[i]] == i, "diagonal block first");
        for (int k = s.h_rowptr[i] + 2; k < s.h_rowptr[i + 1]; k++) check(s.h_col[k - 1] < s.h_col[k], "ascending");
      }
      // the same coordinates through build_from_coords give the same system
      ltb_system c;
      c.natoms = s.natoms;
      c.Z = s.Z;
      c.xyz = s.xyz;
      for (int q = 0; q < 9; q++) c.cell[q] = s.cell[q];
      c.periodic_mask = s.periodic_mask;
      st = ltb::build_from_coords(c);
      check(st.code == 0 && same_system(s, c), "build_from_coords equals build_quartz");
      std::vector<int> order;
      check(ltb::locality_order(s, order).code == 0, "locality_order");
      std::vector<int> seen(s.natoms, 0);
      for (int v : order)
        if (v >= 0 && v < s.natoms) seen[v]++;
      bool perm = (int)order.size() == s.natoms;
      for (int v : seen) perm = perm && v == 1;
      check(perm, "locality_order is a permutation");
      std::vector<int> rp, nj;
      std::vector<double> nd;
      check(ltb::pairs_within(s, 6.0 * ltb::kAngToBohr, rp, nj, nd).code == 0, "pairs_within");
      check((int)rp.size() == s.natoms + 1 && nd.size() == 3 * nj.size(), "pairs_within sizes");
    }
}

void cluster_and_errors() {
  // a 3-atom cluster without a cell
  ltb_system c;
  c.natoms = 3;
  c.Z = {14, 8, 8};
  c.xyz = {0, 0, 0, 3.0, 0, 0, -1.5, 2.6, 0};
  check(ltb::build_from_coords(c).code == 0 && c.h_rowptr[3] == 7, "cluster: 3 diagonal + 4 neighbour blocks");
  // two atoms on top of each other
  ltb_system d;
  d.natoms = 2;
  d.Z = {14, 8};
  d.xyz = {1, 1, 1, 1, 1, 1};
  check(ltb::build_from_coords(d).code != 0, "coincident atoms are refused");
  // more than 20 neighbours: 27 atoms packed in a 1.8 Bohr grid
  ltb_system e;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        e.Z.push_back(8);
        e.xyz.insert(e.xyz.end(), {1.8 * i, 1.8 * j, 1.8 * k});
      }
  e.natoms = 27;
  check(ltb::build_from_coords(e).code == LTB_ERR_NEIGHBORS, "too many neighbours are refused");
  double h[16];
  check(ltb::sk_block(14, 8, 0, 0, 0.001, h).code != 0, "Si-O closer than 0.01 Bohr");
  check(ltb::sk_block(14, 6, 1, 0, 0, h).code != 0, "unknown pair");
}

void matlab_files(const std::string& dir) {
  // a small BSR matrix and a vector, written and read back
  const int rowptr[3] = {0, 2, 3}, col[3] = {0, 1, 1};
  std::vector<double> val(3 * 4);
  for (size_t k = 0; k < val.size(); k++) val[k] = 0.25 * (double)k - 1.0;
  const std::string mp = dir + "/M.m", vp = dir + "/v.m";
  check(ltb_write_matlab_bsr(mp.c_str(), "M", 2, 2, 2, rowptr, col, val.data()) == LTB_OK, "write bsr");
  long long nr = 0, nc = 0, nnz = 0;
  check(ltb_read_matlab_mat(mp.c_str(), &nr, &nc, &nnz, nullptr, nullptr, nullptr) == LTB_OK && nr == 4 && nc == 4,
        "read sizes");
  std::vector<int> r(nnz), cc(nnz);
  std::vector<double> v(nnz);
  check(ltb_read_matlab_mat(mp.c_str(), &nr, &nc, &nnz, r.data(), cc.data(), v.data()) == LTB_OK, "read entries");
  double sum = 0;
  for (double x : v) sum += x;
  double want = 0;
  for (double x : val)
    if (x != 0) want += x;
  check(std::fabs(sum - want) < 1e-12, "bsr round trip");
  long long cap = 1;  // too small a capacity is an error, not an overflow
  check(ltb_read_matlab_mat(mp.c_str(), &nr, &nc, &cap, r.data(), cc.data(), v.data()) != LTB_OK, "capacity check");
  const double vec[4] = {1.5, -2.0, 0.0, 3.25};
  check(ltb_write_matlab_vec(vp.c_str(), "v", 4, vec) == LTB_OK, "write vec");
  long long n = 0;
  check(ltb_read_matlab_vec(vp.c_str(), &n, nullptr) == LTB_OK && n == 4, "vec size");
  double back[4] = {0, 0, 0, 0};
  check(ltb_read_matlab_vec(vp.c_str(), &n, back) == LTB_OK && back[3] == 3.25 && back[1] == -2.0, "vec values");
  check(ltb_write_matlab_vec((dir + "/w.m").c_str(), "not a name", 4, vec) != LTB_OK, "bad MATLAB name");
  // malformed files: empty, cut in the middle of a row, garbage, missing file
  const char* bad[] = {"", "zzz = zeros(2,3);\nzzz = [\n1 1  2.5\n2 ", "hello world\n", "zzz = [\n1 x 3\n];\n"};
  for (int k = 0; k < 4; k++) {
    const std::string p = dir + "/bad" + std::to_string(k) + ".m";
    FILE* f = std::fopen(p.c_str(), "w");
    if (!f) {
      check(false, "cannot create " + p);
      continue;
    }
    std::fputs(bad[k], f);
    std::fclose(f);
    nnz = 0;
    const int st = ltb_read_matlab_mat(p.c_str(), &nr, &nc, &nnz, nullptr, nullptr, nullptr);
    n = 0;
    const int sv = ltb_read_matlab_vec(p.c_str(), &n, nullptr);
    (void)st;
    (void)sv;  // either an error or a consistent (small) size; the sanitizers judge the rest
  }
  check(ltb_read_matlab_vec((dir + "/missing.m").c_str(), &n, nullptr) != LTB_OK, "missing file");
}

// every atom within h hops of the chunk (an independent breadth-first
// distance search) lies in one of the chunk's ranges for hop h
void check_cones(int nat, const std::vector<int>& rp, const std::vector<int>& col, int apc, int max_ranges,
                 const std::string& what) {
  std::vector<std::vector<std::pair<int, int>>> idx;
  std::vector<std::pair<int, int>> ranges;
  ltb::light_cones(nat, rp.data(), col.data(), apc, max_ranges, idx, ranges);
  const int nchunk = (nat + apc - 1) / apc;
  check((int)idx.size() == nchunk, what + ": one cone list per chunk");
  for (int k = 0; k < nchunk && k < (int)idx.size(); k++) {
    std::vector<int> dist(nat, -1), queue;
    for (int a = k * apc; a < std::min((k + 1) * apc, nat); a++) {
      dist[a] = 0;
      queue.push_back(a);
    }
    for (size_t q = 0; q < queue.size(); q++)
      for (int e = rp[queue[q]]; e < rp[queue[q] + 1]; e++)
        if (dist[col[e]] < 0) {
          dist[col[e]] = dist[queue[q]] + 1;
          queue.push_back(col[e]);
        }
    int hmax = 0;
    for (int d : dist) hmax = std::max(hmax, d);
    bool ok = !idx[k].empty();
    for (int h = 0; ok && h <= hmax + 1; h++) {
      const auto& e = idx[k][std::min<size_t>(h, idx[k].size() - 1)];
      ok = e.second >= 1 && e.second <= std::max(1, max_ranges) && e.first + e.second <= (int)ranges.size();
      std::vector<char> cover(nat, 0);
      for (int r = e.first; ok && r < e.first + e.second; r++) {
        const int lo = ranges[r].first, len = ranges[r].second;
        ok = lo >= 0 && lo < nat && len >= 1 && len <= nat;
        for (int i = 0; ok && i < len; i++) cover[(lo + i) % nat] = 1;
      }
      for (int a = 0; ok && a < nat; a++)
        if (dist[a] >= 0 && dist[a] <= h && !cover[a]) ok = false;
    }
    check(ok, what + ": chunk " + std::to_string(k) + " cones cover every atom within reach");
  }
}

// a ring a - 1, a, a + 1 (mod n) over [first, first + n)
void ring(int first, int n, std::vector<int>& rp, std::vector<int>& col) {
  for (int a = 0; a < n; a++) {
    col.push_back(first + a);
    col.push_back(first + (a + 1) % n);
    col.push_back(first + (a + n - 1) % n);
    rp.push_back((int)col.size());
  }
}

void light_cone_paths() {
  ltb_system s;
  ltb::Status st = ltb::build_quartz(3, 3, 3, LTB_LATTICE_CONFIG, 0.05, 1, s);
  check(st.code == 0, "quartz 3x3x3 for the cones");
  for (int apc : {1, 16, 64, 100, 243})
    for (int mr : {1, 2, 32})
      check_cones(s.natoms, s.h_rowptr, s.h_col, apc, mr,
                  "quartz cones apc " + std::to_string(apc) + " ranges " + std::to_string(mr));
  // one ring: the exact cone of chunk 0 at hop h is one cyclic run, its ranges say so
  std::vector<int> rp{0}, col;
  ring(0, 100, rp, col);
  std::vector<std::vector<std::pair<int, int>>> idx;
  std::vector<std::pair<int, int>> ranges;
  ltb::light_cones(100, rp.data(), col.data(), 10, 32, idx, ranges);
  bool exact = idx.size() == 10 && idx[0].size() == 46;
  for (int h = 1; exact && h < 45; h++) {
    const auto& e = idx[0][h];
    exact = e.second == 1 && ranges[e.first].first == 100 - h && ranges[e.first].second == 10 + 2 * h;
  }
  check(exact && ranges[idx[0].back().first] == std::make_pair(0, 100), "ring cones wrap as one range");
  for (int mr : {1, 4, 32}) check_cones(100, rp, col, 10, mr, "ring " + std::to_string(mr));
  // two rings: a cone never leaves its own ring, and its list stops when it stops growing
  std::vector<int> rp2{0}, col2;
  ring(0, 50, rp2, col2);
  ring(50, 50, rp2, col2);
  ltb::light_cones(100, rp2.data(), col2.data(), 25, 32, idx, ranges);
  const auto& last = idx[0].back();
  check(last.second == 1 && ranges[last.first] == std::make_pair(0, 50), "two rings: the cone stays in its ring");
  for (int mr : {1, 3, 32}) check_cones(100, rp2, col2, 25, mr, "two rings " + std::to_string(mr));
  // random symmetric graphs with scattered neighbours (many runs to merge)
  uint64_t x = 12345;
  auto rnd = [&x](int n) {
    x = x * 6364136223846793005ULL + 1442695040888963407ULL;
    return (int)((x >> 33) % (uint64_t)n);
  };
  for (int n : {37, 200}) {
    std::vector<std::vector<int>> adj(n);
    for (int a = 0; a < n; a++) {
      adj[a].push_back(a);
      for (int t = 0; t < 2; t++) {
        const int b = rnd(n);
        if (b != a) {
          adj[a].push_back(b);
          adj[b].push_back(a);
        }
      }
    }
    std::vector<int> rpr{0}, colr;
    for (auto& v : adj) {
      colr.insert(colr.end(), v.begin(), v.end());
      rpr.push_back((int)colr.size());
    }
    for (int apc : {1, 7, 64})
      for (int mr : {1, 2, 5, 32})
        check_cones(n, rpr, colr, apc, mr, "random graph " + std::to_string(n));
  }
}

}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  quartz_paths();
  cluster_and_errors();
  matlab_files(dir);
  light_cone_paths();
  std::vector<double> g(1000);
  ltb::gaussian_fill(3, g.size(), g.data());
  double m = 0;
  for (double x : g) m += x;
  check(std::fabs(m / g.size()) < 0.2, "gaussian stream mean");
  if (g_fail) {
    std::fprintf(stderr, "host_check: %d check(s) failed\n", g_fail);
    return 1;
  }
  std::printf("host_check ok\n");
  return 0;
}
