// MATLAB ASCII (spconvert) files in the layout PETSc's
// PETSC_VIEWER_ASCII_MATLAB viewer writes, which is the on-disk format of every
// dump the reference makes (GAMMA_AB.m, orbit2Atom.m :950-963; deltaH.m,
// netCharge.m, GammaQ_vector.m :1465-1484; H.m, Htotal_scc.m, Htotal0.m
// :1504-1523):
//   %Mat Object: 1 MPI process / %  type: seqaij / % Size = R C / % Nonzeros = K
//   zzz = zeros(K,3); zzz = [ "i j  v" rows, 1-based, row-major ]; NAME = spconvert(zzz);
// and for vectors: %Vec Object:NAME 1 MPI process / %  type: seq / NAME = [ v rows ];
// Values are printed as %.16e, enough to round-trip a double.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ltb.h"

namespace {
thread_local std::string g_io_err;

int io_fail(int code, const std::string& msg) {
  g_io_err = msg;
  return code;
}

bool valid_name(const char* name) {
  if (!name || !*name) return false;
  for (const char* p = name; *p; p++)
    if (!(std::isalnum((unsigned char)*p) || *p == '_')) return false;
  return !std::isdigit((unsigned char)name[0]);
}

// parse "i j v" or "v" lines between "= [" and "];"
int parse(const char* path, bool want_matrix, std::vector<long long>& ri, std::vector<long long>& ci,
          std::vector<double>& v, long long* nr, long long* nc) {
  FILE* f = std::fopen(path, "r");
  if (!f) return io_fail(LTB_ERR_ARG, std::string("cannot open ") + path);
  char line[512];
  bool in = false, is_matrix = false, seen = false;
  *nr = *nc = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (!in) {
      long long a, b;
      if (std::sscanf(line, "%% Size = %lld %lld", &a, &b) == 2) {
        *nr = a;
        *nc = b;
        is_matrix = true;
      }
      if (std::strstr(line, "= [")) {
        in = true;
        seen = true;
      }
      continue;
    }
    if (line[0] == ']') {
      in = false;
      continue;
    }
    char* end = nullptr;
    const double x0 = std::strtod(line, &end);
    if (end == line) continue;
    char* end2 = nullptr;
    const double x1 = std::strtod(end, &end2);
    if (end2 == end) {  // one number: a vector entry
      v.push_back(x0);
      continue;
    }
    char* end3 = nullptr;
    const double x2 = std::strtod(end2, &end3);
    if (end3 == end2) {
      std::fclose(f);
      return io_fail(LTB_ERR_ARG, std::string("malformed entry in ") + path);
    }
    ri.push_back((long long)x0 - 1);
    ci.push_back((long long)x1 - 1);
    v.push_back(x2);
  }
  std::fclose(f);
  if (!seen) return io_fail(LTB_ERR_ARG, std::string("no data block in ") + path);
  if (want_matrix != is_matrix)
    return io_fail(LTB_ERR_ARG, std::string(path) + (want_matrix ? " holds a vector" : " holds a matrix"));
  if (!want_matrix) *nr = (long long)v.size();
  return LTB_OK;
}
}  // namespace

extern "C" {

const char* ltb_io_last_error(void) { return g_io_err.c_str(); }

int ltb_write_matlab_bsr(const char* path, const char* name, int nrows, int ncols, int bs, const int* rowptr,
                         const int* col, const double* val) {
  if (!path || !valid_name(name) || nrows < 0 || ncols < 0 || bs < 1 || !rowptr || (!col && rowptr[nrows]) ||
      (!val && rowptr[nrows]))
    return io_fail(LTB_ERR_ARG, "bad arguments to ltb_write_matlab_bsr");
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(LTB_ERR_ARG, std::string("cannot write ") + path);
  const long long nnz = (long long)rowptr[nrows] * bs * bs;
  std::fprintf(f, "%%Mat Object: 1 MPI process\n%%  type: seqaij\n%% Size = %lld %lld \n%% Nonzeros = %lld \n",
               (long long)nrows * bs, (long long)ncols * bs, nnz);
  std::fprintf(f, "zzz = zeros(%lld,3);\nzzz = [\n", nnz);
  std::vector<int> order;
  for (int i = 0; i < nrows; i++) {
    // AIJ rows list columns in ascending order; the block row may not
    order.clear();
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) order.push_back(k);
    for (size_t a = 1; a < order.size(); a++)
      for (size_t b = a; b > 0 && col[order[b - 1]] > col[order[b]]; b--) std::swap(order[b - 1], order[b]);
    for (int r = 0; r < bs; r++)
      for (int k : order)
        for (int c = 0; c < bs; c++)
          std::fprintf(f, "%lld %lld  %.16e\n", (long long)i * bs + r + 1, (long long)col[k] * bs + c + 1,
                       val[(size_t)k * bs * bs + r * bs + c]);
  }
  std::fprintf(f, "];\n Mat_%s = spconvert(zzz);\n", name);
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? LTB_OK : io_fail(LTB_ERR_ARG, std::string("write error on ") + path);
}

int ltb_write_matlab_vec(const char* path, const char* name, int n, const double* v) {
  if (!path || !valid_name(name) || n < 0 || (n && !v)) return io_fail(LTB_ERR_ARG, "bad arguments to ltb_write_matlab_vec");
  FILE* f = std::fopen(path, "w");
  if (!f) return io_fail(LTB_ERR_ARG, std::string("cannot write ") + path);
  std::fprintf(f, "%%Vec Object:%s 1 MPI process\n%%  type: seq\n%s = [\n", name, name);
  for (int i = 0; i < n; i++) std::fprintf(f, "%.16e\n", v[i]);
  std::fprintf(f, "];\n");
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  return ok ? LTB_OK : io_fail(LTB_ERR_ARG, std::string("write error on ") + path);
}

int ltb_read_matlab_mat(const char* path, long long* nrows, long long* ncols, long long* nnz, int* rows,
                        int* cols, double* vals) {
  if (!path || !nnz) return io_fail(LTB_ERR_ARG, "bad arguments to ltb_read_matlab_mat");
  std::vector<long long> ri, ci;
  std::vector<double> v;
  long long nr, nc;
  int st = parse(path, true, ri, ci, v, &nr, &nc);
  if (st) return st;
  if (nrows) *nrows = nr;
  if (ncols) *ncols = nc;
  const long long cap = *nnz;
  *nnz = (long long)v.size();
  if (!rows && !cols && !vals) return LTB_OK;  // size query
  if (cap < (long long)v.size()) return io_fail(LTB_ERR_ARG, "output arrays too small");
  for (size_t k = 0; k < v.size(); k++) {
    if (ri[k] < 0 || ri[k] >= nr || ci[k] < 0 || ci[k] >= nc)
      return io_fail(LTB_ERR_ARG, std::string("entry out of the declared size in ") + path);
    if (rows) rows[k] = (int)ri[k];
    if (cols) cols[k] = (int)ci[k];
    if (vals) vals[k] = v[k];
  }
  return LTB_OK;
}

int ltb_read_matlab_vec(const char* path, long long* n, double* v) {
  if (!path || !n) return io_fail(LTB_ERR_ARG, "bad arguments to ltb_read_matlab_vec");
  std::vector<long long> ri, ci;
  std::vector<double> x;
  long long nr, nc;
  int st = parse(path, false, ri, ci, x, &nr, &nc);
  if (st) return st;
  const long long cap = *n;
  *n = (long long)x.size();
  if (!v) return LTB_OK;
  if (cap < (long long)x.size()) return io_fail(LTB_ERR_ARG, "output array too small");
  for (size_t k = 0; k < x.size(); k++) v[k] = x[k];
  return LTB_OK;
}

}  // extern "C"
