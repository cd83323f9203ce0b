"""ctypes front end of the CPU FP64 oracle (oracle.c), plus two numpy
restatements for the stochastic mode.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Pins of the numpy functions (the C functions' pins are in oracle.c):
* kpm_dos (runKPM :1634-1701): its Jackson kernel is `jackson`, pinned by
  the reference's gnsfile.txt; on exact trace moments of a 72-atom cell the DOS
  integrates to the orbital count, mu falls in the HOMO-LUMO gap of the dense
  eigenvalues and the narrowed bounds sit at the extreme eigenvalues
  (tests/test_stochastic_host.py).
* gaussian_stream: reproduces the atom displacements that orc_quartz draws
  (displacement / sigma, 1e-12).
* System.scc_two_level (the two-level SCC of ltb_scc_run with M_lo > 0, a
  loop over orc_density): converges to the fixed point of the plain orc_scc
  at the full order, from the neutral start and from a perturbed one, and a
  planted error in its defect (dropped, or taken from the wrong order) moves
  the answer or stalls it (tests/test_oracle_math.py).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
MAXN = 20
MODE_CONFIG = 0
MODE_REFERENCE = 1
HARTREE_EV = 27.211386

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _setup(_lib)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)


class KpeParams(C.Structure):
    _fields_ = [("M", C.c_int), ("kT", C.c_double), ("K", C.c_int), ("Ne", C.c_double),
                ("bisect_iters", C.c_int)]


class KpeResult(C.Structure):
    _fields_ = [("mu", C.c_double), ("mu_ha", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("emin", C.c_double), ("emax", C.c_double), ("nelec", C.c_double),
                ("eband", C.c_double)]


def _setup(L):
    L.orc_quartz.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, PD, PI, PD]
    L.orc_neighbors.argtypes = [C.c_int, PD, PD, C.c_int, PI, PI, PD]
    L.orc_sk_block.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, PD]
    L.orc_build_h.argtypes = [C.c_int, PI, PI, PI, PD, C.c_int, PI, PI, PD]
    L.orc_build_gamma.argtypes = [C.c_int, PI, PI, PD, C.c_double, C.c_int, PI, PI, PD]
    L.orc_gershgorin.argtypes = [C.c_int, PI, PI, PD, PD, PD, PD]
    L.orc_gmM.argtypes = [C.c_int, PD]
    L.orc_gmM.restype = None
    L.orc_jackson.argtypes = [C.c_int, PD]
    L.orc_jackson.restype = None
    L.orc_cheb_nodes.argtypes = [C.c_int, PD]
    L.orc_cheb_nodes.restype = None
    L.orc_step_coeffs.argtypes = [C.c_int, C.c_double, PD]
    L.orc_step_coeffs.restype = None
    L.orc_fd_coeffs.argtypes = [C.c_int, C.c_double, C.c_double, C.c_int, PD]
    L.orc_fd_coeffs.restype = None
    L.orc_rescaled_csr.argtypes = [C.c_int, PI, PI, PD, PD, C.c_double, C.c_double, PI, PI, PD]
    L.orc_kpe_diag_moments.argtypes = [C.c_int, PI, PI, PD, C.c_int, C.c_int, C.c_int, PD, C.c_int]
    L.orc_kpe_diag_moments_cols.argtypes = [C.c_int, PI, PI, PD, C.c_int, C.c_int, PI, PD, C.c_int]
    L.orc_kpe_apply.argtypes = [C.c_int, PI, PI, PD, C.c_int, PD, PD, PD]
    L.orc_count.argtypes = [C.c_int, PD, PD]
    L.orc_count.restype = C.c_double
    L.orc_density.argtypes = [C.c_int, PI, PI, PD, PD, C.POINTER(KpeParams), PD, PD,
                              C.POINTER(KpeResult), C.c_int]
    L.orc_scc.argtypes = [C.c_int, PI, PI, PD, PI, PI, PD, PD, C.POINTER(KpeParams), C.c_double,
                          C.c_double, C.c_int, PD, PI, PD, C.POINTER(KpeResult), C.c_int]
    L.orc_sp2_dense.argtypes = [C.c_int, PD, C.c_double, C.c_int, C.c_double, PD]


class OracleError(RuntimeError):
    pass


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what} failed with status {st}")


def quartz(L, M, N, mode=MODE_CONFIG, sigma_ang=0.0, seed=1):
    nat = 9 * L * M * N
    xyz = np.zeros(3 * nat)
    Z = np.zeros(nat, np.int32)
    cell = np.zeros(9)
    _check(lib().orc_quartz(L, M, N, mode, sigma_ang, seed, _p(xyz, C.c_double), _p(Z, C.c_int),
                            _p(cell, C.c_double)), "orc_quartz")
    return xyz.reshape(nat, 3), Z, cell.reshape(3, 3)


class System:
    """Geometry + neighbour list + H (BSR 4x4) + Gamma (CSR), all FP64."""

    def __init__(self, xyz, Z, cell, periodic_mask=7, Us=1.0):
        self.xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        self.Z = np.ascontiguousarray(Z, np.int32)
        self.cell = np.ascontiguousarray(cell, np.float64).reshape(3, 3)
        nat = self.nat = len(self.Z)
        self.norb = 4 * nat
        nn = np.zeros(nat, np.int32)
        nbr = np.zeros(nat * MAXN, np.int32)
        nd = np.zeros(nat * MAXN * 3)
        _check(lib().orc_neighbors(nat, _p(self.xyz, C.c_double), _p(self.cell, C.c_double),
                                   periodic_mask, _p(nn, C.c_int), _p(nbr, C.c_int),
                                   _p(nd, C.c_double)), "orc_neighbors")
        self.nn, self.nbr, self.nd = nn, nbr, nd
        cap = nat * (MAXN + 1)
        rp = np.zeros(nat + 1, np.int32)
        col = np.zeros(cap, np.int32)
        val = np.zeros(cap * 16)
        nz = lib().orc_build_h(nat, _p(self.Z, C.c_int), _p(nn, C.c_int), _p(nbr, C.c_int),
                               _p(nd, C.c_double), cap, _p(rp, C.c_int), _p(col, C.c_int),
                               _p(val, C.c_double))
        if nz < 0:
            raise OracleError(f"orc_build_h failed with status {-nz}")
        self.h_rowptr, self.h_col, self.h_val = rp, col[:nz].copy(), val[:nz * 16].copy()
        grp = np.zeros(nat + 1, np.int32)
        gcol = np.zeros(cap, np.int32)
        gval = np.zeros(cap)
        nz = lib().orc_build_gamma(nat, _p(nn, C.c_int), _p(nbr, C.c_int), _p(nd, C.c_double), Us,
                                   cap, _p(grp, C.c_int), _p(gcol, C.c_int), _p(gval, C.c_double))
        if nz < 0:
            raise OracleError(f"orc_build_gamma failed with status {-nz}")
        self.g_rowptr, self.g_col, self.g_val = grp, gcol[:nz].copy(), gval[:nz].copy()
        self.nelectrons = float(np.sum(np.where(self.Z == 14, 4.0, 6.0)))
        self.p0 = np.where(self.Z == 14, 2.0, 3.0)  # neutral per-spin populations

    @classmethod
    def quartz(cls, L, M, N, mode=MODE_CONFIG, sigma_ang=0.0, seed=1):
        xyz, Z, cell = quartz(L, M, N, mode, sigma_ang, seed)
        return cls(xyz, Z, cell, periodic_mask=(3 if mode == MODE_REFERENCE else 7))

    @property
    def nnzb(self):
        return len(self.h_col)

    def dense_h(self, shift=None):
        n = self.norb
        H = np.zeros((n, n))
        v = self.h_val.reshape(-1, 4, 4)
        for i in range(self.nat):
            for k in range(self.h_rowptr[i], self.h_rowptr[i + 1]):
                j = self.h_col[k]
                H[4 * i:4 * i + 4, 4 * j:4 * j + 4] = v[k]
        if shift is not None:
            H[np.diag_indices(n)] += shift
        return H

    def set_gamma_onsite(self, gamma_aa):
        """Gamma_AA per atom (the diagonal entries of the CSR rows), in place."""
        for i in range(self.nat):
            for k in range(self.g_rowptr[i], self.g_rowptr[i + 1]):
                if self.g_col[k] == i:
                    self.g_val[k] = gamma_aa[i]

    def dense_gamma(self):
        G = np.zeros((self.nat, self.nat))
        for i in range(self.nat):
            for k in range(self.g_rowptr[i], self.g_rowptr[i + 1]):
                G[i, self.g_col[k]] = self.g_val[k]
        return G

    def gershgorin(self, shift=None):
        lo, hi = C.c_double(), C.c_double()
        sp = _p(np.ascontiguousarray(shift, np.float64), C.c_double) if shift is not None else None
        _check(lib().orc_gershgorin(self.nat, _p(self.h_rowptr, C.c_int), _p(self.h_col, C.c_int),
                                    _p(self.h_val, C.c_double), sp, C.byref(lo), C.byref(hi)),
               "orc_gershgorin")
        return lo.value, hi.value

    def rescaled_csr(self, a, b, shift=None):
        n = self.norb
        nnz = self.nnzb * 16
        crow = np.zeros(n + 1, np.int32)
        ccol = np.zeros(nnz, np.int32)
        cval = np.zeros(nnz)
        sp = None
        if shift is not None:
            shift = np.ascontiguousarray(shift, np.float64)
            sp = _p(shift, C.c_double)
        lib().orc_rescaled_csr(self.nat, _p(self.h_rowptr, C.c_int), _p(self.h_col, C.c_int),
                               _p(self.h_val, C.c_double), sp, a, b, _p(crow, C.c_int),
                               _p(ccol, C.c_int), _p(cval, C.c_double))
        return crow, ccol, cval

    def diag_moments(self, a, b, nm, j0=0, j1=None, shift=None, nthreads=0):
        j1 = self.norb if j1 is None else j1
        crow, ccol, cval = self.rescaled_csr(a, b, shift)
        out = np.zeros((j1 - j0) * nm)
        _check(lib().orc_kpe_diag_moments(self.norb, _p(crow, C.c_int), _p(ccol, C.c_int),
                                          _p(cval, C.c_double), nm, j0, j1, _p(out, C.c_double),
                                          nthreads), "orc_kpe_diag_moments")
        return out.reshape(j1 - j0, nm)

    def diag_moments_at(self, a, b, nm, cols, shift=None, nthreads=0):
        """[T_n]_jj, n < nm, for the listed orbitals j: [len(cols), nm]"""
        cols = np.ascontiguousarray(cols, np.int32)
        crow, ccol, cval = self.rescaled_csr(a, b, shift)
        out = np.zeros(len(cols) * nm)
        _check(lib().orc_kpe_diag_moments_cols(self.norb, _p(crow, C.c_int), _p(ccol, C.c_int),
                                               _p(cval, C.c_double), nm, len(cols), _p(cols, C.c_int),
                                               _p(out, C.c_double), nthreads), "orc_kpe_diag_moments_cols")
        return out.reshape(len(cols), nm)

    def kpe_apply(self, a, b, w, x, shift=None):
        crow, ccol, cval = self.rescaled_csr(a, b, shift)
        w = np.ascontiguousarray(w, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.norb)
        _check(lib().orc_kpe_apply(self.norb, _p(crow, C.c_int), _p(ccol, C.c_int),
                                   _p(cval, C.c_double), len(w), _p(w, C.c_double),
                                   _p(x, C.c_double), _p(out, C.c_double)), "orc_kpe_apply")
        return out

    def kpe_params(self, M, kT_ev=0.05, K=8192, bisect_iters=100):
        kT = kT_ev / HARTREE_EV if kT_ev > 0 else 0.0
        return KpeParams(M, kT, K, self.nelectrons, bisect_iters)

    def density(self, M, shift=None, kT_ev=0.05, K=8192, nthreads=0, want_moments=False):
        p = self.kpe_params(M, kT_ev, K)
        pop = np.zeros(self.nat)
        res = KpeResult()
        mom = np.zeros(self.nat * (M + 1)) if want_moments else None
        sp = None
        if shift is not None:
            shift = np.ascontiguousarray(shift, np.float64)
            sp = _p(shift, C.c_double)
        _check(lib().orc_density(self.nat, _p(self.h_rowptr, C.c_int), _p(self.h_col, C.c_int),
                                 _p(self.h_val, C.c_double), sp, C.byref(p), _p(pop, C.c_double),
                                 _p(mom, C.c_double) if mom is not None else None, C.byref(res),
                                 nthreads), "orc_density")
        out = dict(pop=pop, mu=res.mu, mu_ha=res.mu_ha, a=res.a, b=res.b, emin=res.emin,
                   emax=res.emax, nelec=res.nelec, eband=res.eband)
        if mom is not None:
            out["moments"] = mom.reshape(self.nat, M + 1)
        return out

    def scc(self, M, guess=None, kT_ev=0.05, K=8192, mix=0.2, tol=1e-6, maxit=200, nthreads=0):
        p = self.kpe_params(M, kT_ev, K)
        pio = np.ascontiguousarray(self.p0 if guess is None else guess, np.float64).copy()
        its = C.c_int()
        hist = np.zeros(maxit)
        res = KpeResult()
        st = lib().orc_scc(self.nat, _p(self.h_rowptr, C.c_int), _p(self.h_col, C.c_int),
                           _p(self.h_val, C.c_double), _p(self.g_rowptr, C.c_int),
                           _p(self.g_col, C.c_int), _p(self.g_val, C.c_double),
                           _p(self.p0, C.c_double), C.byref(p), mix, tol, maxit,
                           _p(pio, C.c_double), C.byref(its), _p(hist, C.c_double),
                           C.byref(res), nthreads)
        return dict(status=st, pop=pio, q=2 * pio, iters=its.value, hist=hist[:its.value],
                    mu=res.mu, mu_ha=res.mu_ha, a=res.a, b=res.b, eband=res.eband,
                    nelec=res.nelec)

    def scc_two_level(self, M, M_lo, **kw):
        return _scc_two_level(self, M, M_lo, **kw)


def _scc_two_level(o, M, M_lo, guess=None, kT_ev=0.05, K=8192, mix=0.2, tol=1e-6, lo_tol=0.0, maxit=2000,
                   nthreads=0, defect_order=None, max_phase=400):
    """Two-level SCC in FP64 (the control flow of ltb_scc_run with M_lo > 0):
    a full-order iteration evaluates F_M and F_M_lo at q_in, keeps the defect
    d = F_M - F_M_lo and mixes with F_M; low-order iterations mix with
    F_M_lo + d until their |dq| < lo_tol or < 1e-3 of the last full-order
    |dq|, or their smallest |dq| has not fallen by 10% for 4 iterations, or
    max_phase of them; convergence only on
    a full-order iteration with |dq| < tol.  defect_order (test hook) takes
    the defect's low-order term at another order."""
    lo_tol = lo_tol if lo_tol > 0 else 0.1 * tol
    G = o.dense_gamma()
    p0 = o.p0
    pin = np.array(p0 if guess is None else guess, np.float64)
    d = np.zeros(o.nat)
    full, phase, n_lo, hist, full_hist = True, 0, 0, [], []
    best, stall, dq_full = 0.0, 0, 0.0
    status = 1
    for it in range(maxit):
        shift = np.repeat(G @ (pin - p0), 4)
        if full:
            lo = o.density(defect_order or M_lo, shift=shift, kT_ev=kT_ev, K=K, nthreads=nthreads)
            r = o.density(M, shift=shift, kT_ev=kT_ev, K=K, nthreads=nthreads)
            pout = r["pop"]
            d = pout - lo["pop"]
        else:
            r = o.density(M_lo, shift=shift, kT_ev=kT_ev, K=K, nthreads=nthreads)
            pout = r["pop"] + d
            n_lo += 1
            phase += 1
        dq = 2 * np.abs(pout - pin).max()
        hist.append(dq)
        if full:
            full_hist.append(dq)
        if full and dq < tol:
            status = 0
            break
        pin = pin + mix * (pout - pin)
        if full:
            full, phase, dq_full = False, 0, dq
        else:
            if phase == 1 or dq < 0.9 * best:
                best, stall = dq, 0
            else:
                stall += 1
            if dq < lo_tol or dq < 1e-3 * dq_full or stall >= 4 or phase >= max_phase:
                full = True
    return dict(status=status, pop=pout, q=2 * pout, iters=len(hist), lo_iters=n_lo, hist=np.array(hist),
                full_hist=np.array(full_hist), mu_ha=r["mu_ha"], eband=r["eband"])


def gmM(M):
    g = np.zeros(M)
    lib().orc_gmM(M, _p(g, C.c_double))
    return g


def jackson(N):
    g = np.zeros(N)
    lib().orc_jackson(N, _p(g, C.c_double))
    return g


def cheb_nodes(n):
    x = np.zeros(n)
    lib().orc_cheb_nodes(n, _p(x, C.c_double))
    return x


def step_coeffs(M, mu):
    c = np.zeros(M)
    lib().orc_step_coeffs(M, mu, _p(c, C.c_double))
    return c


def fd_coeffs(M, mu, kT, K=8192):
    c = np.zeros(M)
    lib().orc_fd_coeffs(M, mu, kT, K, _p(c, C.c_double))
    return c


def kpe_weights(c, g):
    w = c * g
    w[0] = c[0] / 2.0
    return w


def sk_block(zi, zj, d):
    h = np.zeros(16)
    _check(lib().orc_sk_block(zi, zj, float(d[0]), float(d[1]), float(d[2]), _p(h, C.c_double)),
           "orc_sk_block")
    return h.reshape(4, 4)


def sp2_dense(X, nocc, maxit=100, idem_tol=1e-12):
    X = np.ascontiguousarray(X, np.float64).copy()
    idem = C.c_double()
    it = lib().orc_sp2_dense(X.shape[0], _p(X, C.c_double), nocc, maxit, idem_tol, C.byref(idem))
    if it < 0:
        raise OracleError("orc_sp2_dense failed")
    return X, it, idem.value


def kpm_dos(moments, a, b, nsize, nelectrons, ntilde=None):
    """runKPM's DOS, mu and bound tightening (withSP2_deltaH.cpp:1634-1701),
    literally: Jackson g_n with N = len(moments), DOS[k] = (g_0 mu_0 + 2 sum_j
    g_j mu_j cos(j acos x_k)) / (pi sqrt(1 - x_k^2)) at x_k = cos(pi (k + 1/2) /
    ntilde), trapezoid area from the bottom, mu at area fraction Ne / (2 nsize),
    bounds where DOS / nsize >= 0.001."""
    mom = np.asarray(moments, np.float64)
    n = len(mom)
    nt = 2 * n if ntilde is None else ntilde
    g = jackson(n)
    x = np.cos(np.pi * (np.arange(nt) + 0.5) / nt)
    j = np.arange(1, n)
    dos = np.array([(g[0] * mom[0] + 2 * np.sum(g[1:] * mom[1:] * np.cos(j * np.arccos(xv))))
                    / (np.pi * np.sqrt(1.0 - xv * xv)) for xv in x])
    E = x * a + b
    seg = (x[:-1] - x[1:]) * (dos[1:] + dos[:-1]) / 2   # seg[i-1] = trapezoid between x[i-1] and x[i]
    total = seg[::-1].sum()
    frac = nelectrons * total / (2.0 * nsize)
    acc, imu = 0.0, 0
    for i in range(nt - 1, 0, -1):
        acc += seg[i - 1]
        if acc >= frac:
            imu = i
            break
    lo, hi = -10.0, 10.0
    for i in range(nt - 1, 0, -1):
        if dos[i] / nsize >= 0.001:
            lo = E[i]
            break
    for i in range(nt):
        if dos[i] / nsize >= 0.001:
            hi = E[i]
            break
    return dict(energies=E, dos=dos, mu=E[imu], emin=lo, emax=hi, area=total)


def gaussian_stream(seed, n):
    """n standard normals of the SplitMix64 -> Box-Muller stream (uniforms
    ((z >> 11) + 1) 2^-53, pairs r cos t then r sin t)."""
    mask = (1 << 64) - 1
    s = seed & mask
    out = np.empty(n)
    k = 0
    while k < n:
        u = []
        for _ in range(2):
            s = (s + 0x9E3779B97F4A7C15) & mask
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            z ^= z >> 31
            u.append(((z >> 11) + 1) * 2.0 ** -53)
        r = np.sqrt(-2.0 * np.log(u[0]))
        t = 2.0 * np.pi * u[1]
        out[k] = r * np.cos(t)
        if k + 1 < n:
            out[k + 1] = r * np.sin(t)
        k += 2
    return out
