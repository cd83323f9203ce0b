// Device context shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/ltb.h"

// scalars that stay on the device so one SCC iteration needs no host round trip
struct DevScalars {
  double emin, emax;   // Gershgorin bounds of H_total
  double a, b;         // rescale (H - b)/a
  double mu;           // scaled chemical potential
  double nelec;        // 2 * sum w_m mu_m
  double eband;        // 2 Tr(P H_total)
  double dq;           // |q_out - q_in|_inf (e)
  double eband_fp;     // eband + 2 pop . Gamma (p_out - p_in): fixed-point estimate
  int nonfinite;       // set by reductions that met NaN/Inf
  int pad;
};

struct ltb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int cu_count = 0;
  // system
  int nat = 0, norb = 0, nnzb = 0, gnnz = 0;
  double nelectrons = 0;
  int *d_rowptr = nullptr, *d_col = nullptr, *d_blkrow = nullptr;
  double* d_h0 = nullptr;       // [nnzb][16]
  double* d_ht64 = nullptr;     // rescaled
  float* d_ht32 = nullptr;
  double* d_offabs = nullptr;   // per orbital sum |H0_ij|, j != i
  double* d_diag0 = nullptr;    // per orbital H0_ii
  int *d_grow = nullptr, *d_gcol = nullptr;
  double* d_gval = nullptr;
  double *d_p0 = nullptr, *d_pin = nullptr, *d_pout = nullptr, *d_v = nullptr;
  DevScalars* d_sc = nullptr;
  double* d_red = nullptr;      // reduction scratch [4096]
  // KPE buffers
  void *d_xa = nullptr, *d_xb = nullptr;
  size_t xabytes = 0, xbbytes = 0;
  double* d_part = nullptr;
  size_t partbytes = 0;
  double* d_mu1 = nullptr;      // per column [T_1]_jj
  int* d_visit = nullptr;       // sweep visiting order (ltb::locality_order)
  size_t mu1bytes = 0;
  double* d_mom = nullptr;      // [M+1][nat]
  double* d_momtot = nullptr;   // [M+1]
  int mom_M = -1;
  bool keep_orbital = false;    // also keep per-orbital moments [M+1][norb]
  double* d_momorb = nullptr;
  size_t momorbbytes = 0;
  int momorb_M = -1;
  size_t mombytes = 0;
  // Fermi / density
  double *d_W = nullptr, *d_X = nullptr, *d_g = nullptr, *d_c = nullptr, *d_w = nullptr;
  int fd_K = 0, fd_M = 0;
  // the same for the low order of the two-level SCC (swapped in while its
  // density runs, so alternating orders reallocates nothing)
  double *lo_W = nullptr, *lo_X = nullptr, *lo_g = nullptr, *lo_w = nullptr;
  int lo_K = 0, lo_M = 0;
  // two-level SCC: low-order populations at q_in and the defect F_M - F_M_lo [nat]
  double *d_plo = nullptr, *d_defect = nullptr;
  // tiling of the Chebyshev kernel per recursion precision, [0] FP32 and
  // [1] FP64 (0 / -1 fields = auto)
  ltb_kpe_tiling tiling[2] = {{0, 0, 0, 0, 0, -1, 0, 0}, {0, 0, 0, 0, 0, -1, 0, 0}};
  // statistics of the last ltb_kpe_moments
  int stat_launches = 0;
  double stat_seconds = 0, stat_bytes = 0;
  bool timing = false;
  std::vector<hipEvent_t> events;
  // SP2 dense buffers
  void *d_sa = nullptr, *d_sb = nullptr;
  size_t sbytes = 0;
  // device address of the allocation-free entry point's scratch array on this
  // context's device (not owned: the array lives as long as the process)
  void* d_scratch = nullptr;
  // light cones of the unit-vector chunks (sweep schedule), per chunk width
  // ld: for chunk k and hop count h, a few cyclic atom ranges {lo, len} that
  // hold every atom within h hops of the chunk's atoms; built from a host
  // copy of the block pattern, dropped with the system
  struct ConeTable {
    std::vector<std::vector<std::pair<int, int>>> idx;  // [chunk][hop]: first range, number of ranges
    std::vector<std::pair<int, int>> ranges;            // {lo, len} of every (chunk, hop), host copy
    int* d_ranges = nullptr;                            // the same on the device, 2 ints per range
  };
  bool light_cone = true;
  std::map<int, ConeTable> cones;
  std::vector<int> h_rowptr, h_col;
};

int set_err(ltb_ctx* c, int code, const std::string& msg);

// End of a low-order phase of the two-level SCC (ltb_scc_run and the batch
// alike): after kTwoLevelMaxPhase iterations even if its |dq| is still above
// lo_tol (linear mixing 0.2 needs ~70 from a neutral start); once its
// smallest |dq| has not fallen by 10% for kTwoLevelStall iterations (in FP32
// the corrected map's |dq| stops at the rounding level of the recursion,
// ~1.5e-7 e at 15,552 atoms, while mixing alone shrinks it by 1 - mix per
// iteration); or below kTwoLevelPhaseRel times the last full-order |dq|: the
// next full-order iteration is off by about rho |dq_full| anyway (rho ~ 0.015
// at M_lo = 32, the error of the corrected map), so resolving the phase's own
// fixed point more finely than that buys nothing.
constexpr int kTwoLevelMaxPhase = 400, kTwoLevelStall = 4;
constexpr double kTwoLevelPhaseRel = 1e-3;

// instantiated graphs of one SCC run, destroyed on every exit path
template <int N>
struct GraphExecs {
  hipGraphExec_t e[N] = {};
  GraphExecs() = default;
  GraphExecs(const GraphExecs&) = delete;
  GraphExecs& operator=(const GraphExecs&) = delete;
  ~GraphExecs() {
    for (hipGraphExec_t x : e)
      if (x) (void)hipGraphExecDestroy(x);
  }
};

#define LTB_HIP(ctx, call)                                                              \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return set_err((ctx), LTB_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// helpers implemented in ltb_capi.hip, used by sp2.hip and assemble.hip
void release_system(ltb_ctx* c);
int alloc_system(ltb_ctx* c, int nat, int nnzb, int gnnz);
int ensure_buffer(ltb_ctx* c, void** p, size_t* cap, size_t bytes);
int kpe_moments_impl(ltb_ctx* c, int M, int precision);
// size the recursion, partial and moment buffers for order M in this precision
int kpe_reserve(ltb_ctx* c, int M, int precision);
// drop the light-cone tables (with the system or the context)
void free_cones(ltb_ctx* c);
// moments r_c^T T_m r_c of explicit vectors already in d_xa (row-major [norb][ld])
int kpe_block_moments_impl(ltb_ctx* c, int M, int nvec, int precision, const double* d_r2, double* d_momcol,
                           int ld);
int kpe_block_width(ltb_ctx* c, int precision, int nvec);
// getKPE on a device block: a holds t_0 (norb x ld, row-major, overwritten), b
// is scratch of the same size, theta (FP64, norb x ld) receives the result
int kpe_apply_device(ltb_ctx* c, int M, const double* w, int ld, void* a, void* b, double* theta, int precision);
int kpe_time_steps(ltb_ctx* c, int precision, int steps, double* seconds, double* bytes);
int kpe_resolved_tiling(ltb_ctx* c, int precision, ltb_kpe_tiling* out);
// whole exact-trace recursion of order M at chunk width `cols` (light cones as
// set), device seconds of the fastest of `reps` runs after a warm-up
int kpe_time_recursion(ltb_ctx* c, int precision, int M, int cols, int reps, double* seconds);
typedef ltb_kpe_tiling KpeTiling;
int noalloc_scratch_init(ltb_ctx* c);
int sp2_dense_impl(ltb_ctx* c, double Ne, int max_iter, int fp64_tail, double* diagP, double* pop,
                   ltb_sp2_result* res);
