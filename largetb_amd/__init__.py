"""largetb_amd: MI355X-native density-matrix path (KPE / SP2 / SCC) of
pghale/largeTB behind a C ABI (include/ltb.h).

This module is a ctypes front end of the in-tree ``libltb.so``.  It raises if
the library is missing: there is no CPU fallback for the device calls.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libltb.so")

LTB_OK = 0
LTB_ERR_ARG = 1
LTB_ERR_GEOMETRY = 2
LTB_ERR_NEIGHBORS = 3
LTB_ERR_NONFINITE = 4
LTB_NOT_CONVERGED = 5
LTB_ERR_DEVICE = 6
LTB_ERR_STATE = 7
LTB_ERR_NOMEM = 8
LATTICE_CONFIG = 0
LATTICE_REFERENCE = 1
FP64 = 0
FP32_ACC64 = 1
HARTREE_EV = 27.211386

_lib = None


class LtbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"status {status}: {msg}")
        self.status = status


class QuartzDesc(C.Structure):
    _fields_ = [("L", C.c_int), ("M", C.c_int), ("N", C.c_int), ("lattice", C.c_int),
                ("sigma_ang", C.c_double), ("seed", C.c_ulonglong)]


class FermiParams(C.Structure):
    _fields_ = [("M", C.c_int), ("kT", C.c_double), ("K", C.c_int), ("Ne", C.c_double)]


class DensityResult(C.Structure):
    _fields_ = [("mu", C.c_double), ("mu_ha", C.c_double), ("a", C.c_double), ("b", C.c_double),
                ("nelec", C.c_double), ("eband", C.c_double)]


class KpeTiling(C.Structure):
    _fields_ = [("chunk_cols", C.c_int), ("cols_per_lane", C.c_int), ("schedule", C.c_int),
                ("blocks_per_cu", C.c_int), ("atoms_per_wave", C.c_int), ("nontemporal", C.c_int),
                ("order", C.c_int), ("lo_chunk_cols", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SCHED_AUTO, SCHED_BLOCKED, SCHED_SWEEP = 0, 1, 2


# Hubbard parameters the reference declares in its on-site Slater-Koster
# blocks but never uses (withSP2_deltaH.cpp:261 Si Ud, :303 O Up), Hartree
HUBBARD_U = {14: 0.247609, 8: 0.4954}


def hubbard_u(Z):
    """Gamma_AA = +U per atom for ltb_set_gamma_onsite (the +U variant)."""
    return np.array([HUBBARD_U[int(z)] for z in Z], np.float64)


class SccParams(C.Structure):
    _fields_ = [("M", C.c_int), ("kT", C.c_double), ("K", C.c_int), ("mix", C.c_double),
                ("tol", C.c_double), ("max_iter", C.c_int), ("precision", C.c_int),
                ("switch_tol", C.c_double), ("use_graph", C.c_int), ("verbose", C.c_int),
                ("M_lo", C.c_int), ("lo_tol", C.c_double)]


class SccResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("fp64_iterations", C.c_int), ("dq", C.c_double),
                ("mu_ha", C.c_double), ("eband", C.c_double), ("eband_fp", C.c_double),
                ("graph_launches", C.c_int), ("seconds", C.c_double), ("lo_iterations", C.c_int)]


class Sp2Result(C.Structure):
    _fields_ = [("iterations", C.c_int), ("fp64_iterations", C.c_int), ("trace", C.c_double),
                ("idem", C.c_double), ("gemm_seconds", C.c_double), ("gemm_flops", C.c_double)]


PI = C.POINTER(C.c_int)
PD = C.POINTER(C.c_double)
VP = C.c_void_p

# every symbol declared in include/ltb.h with its ctypes signature
SIGNATURES = {
    "ltb_system_quartz": (C.c_int, [C.POINTER(QuartzDesc), C.POINTER(VP)]),
    "ltb_system_create": (C.c_int, [C.c_int, PI, PD, PD, C.c_int, C.POINTER(VP)]),
    "ltb_system_destroy": (None, [VP]),
    "ltb_system_last_error": (C.c_char_p, []),
    "ltb_system_sizes": (C.c_int, [VP, PI, PI, PI, PD]),
    "ltb_system_coords": (C.c_int, [VP, PI, PD]),
    "ltb_system_hamiltonian": (C.c_int, [VP, PI, PI, PD]),
    "ltb_system_gamma": (C.c_int, [VP, PI, PI, PD]),
    "ltb_system_neutral": (C.c_int, [VP, PD]),
    "ltb_write_matlab_bsr": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, PI, PI, PD]),
    "ltb_write_matlab_vec": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, PD]),
    "ltb_read_matlab_mat": (C.c_int, [C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                      C.POINTER(C.c_longlong), PI, PI, PD]),
    "ltb_read_matlab_vec": (C.c_int, [C.c_char_p, C.POINTER(C.c_longlong), PD]),
    "ltb_io_last_error": (C.c_char_p, []),
    "ltb_create": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "ltb_destroy": (None, [VP]),
    "ltb_last_error": (C.c_char_p, [VP]),
    "ltb_load_system": (C.c_int, [VP, VP]),
    "ltb_load_coords": (C.c_int, [VP, C.c_int, PI, PD, PD, C.c_int, PD]),
    "ltb_get_sizes": (C.c_int, [VP, PI, PI, PI, PD]),
    "ltb_get_hamiltonian": (C.c_int, [VP, PI, PI, PD]),
    "ltb_get_gamma": (C.c_int, [VP, PI, PI, PD]),
    "ltb_set_populations": (C.c_int, [VP, PD]),
    "ltb_set_neutral": (C.c_int, [VP, PD]),
    "ltb_get_shift": (C.c_int, [VP, PD]),
    "ltb_get_mixed_populations": (C.c_int, [VP, PD]),
    "ltb_gershgorin": (C.c_int, [VP, PD, PD]),
    "ltb_rescale": (C.c_int, [VP, C.c_double, C.c_double]),
    "ltb_kpe_moments": (C.c_int, [VP, C.c_int, C.c_int, PD, PD]),
    "ltb_density": (C.c_int, [VP, C.POINTER(FermiParams), PD, C.POINTER(DensityResult)]),
    "ltb_scc_run": (C.c_int, [VP, C.POINTER(SccParams), PD, PD, PD, C.POINTER(SccResult)]),
    "ltb_kpe_apply": (C.c_int, [VP, C.c_int, PD, C.c_int, PD, PD, C.c_int]),
    "ltb_kpe_apply_noalloc": (C.c_int, [VP, C.c_int, PD, C.c_int, VP, VP, C.c_int]),
    "ltb_kpe_block_moments": (C.c_int, [VP, C.c_int, C.c_int, PD, C.c_int, PD, PD]),
    "ltb_random_vectors": (C.c_int, [C.c_ulonglong, C.c_int, C.c_int, PD]),
    "ltb_kpe_stochastic": (C.c_int, [VP, C.c_int, C.c_int, C.c_ulonglong, C.c_int, PD]),
    "ltb_density_stochastic": (C.c_int, [VP, C.c_int, PD, C.c_int, C.c_ulonglong, C.c_int, PD]),
    "ltb_trace_stochastic": (C.c_int, [VP, C.c_int, PD, C.c_int, C.c_ulonglong, C.c_int, PD]),
    "ltb_kpm_dos": (C.c_int, [PD, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_double, PD, PD, PD, PD,
                              PD]),
    "ltb_write_dos": (C.c_int, [C.c_char_p, C.c_int, PD, PD]),
    "ltb_sp2_dense": (C.c_int, [VP, C.c_double, C.c_int, C.c_int, PD, PD, C.POINTER(Sp2Result)]),
    "ltb_kpe_stats": (C.c_int, [VP, PI, PD, PD]),
    "ltb_set_kpe_orbital_moments": (C.c_int, [VP, C.c_int]),
    "ltb_set_kpe_light_cone": (C.c_int, [VP, C.c_int]),
    "ltb_set_gamma_onsite": (C.c_int, [VP, PD]),
    "ltb_density_diag": (C.c_int, [VP, PD]),
    "ltb_get_orbital_moments": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, PD]),
    "ltb_kpe_weights": (C.c_int, [VP, C.c_int, PD]),
    "ltb_kpe_weights_at": (C.c_int, [VP, C.c_int, C.c_double, C.c_double, C.c_int, PD]),
    "ltb_set_kpe_timing": (C.c_int, [VP, C.c_int]),
    "ltb_set_kpe_tiling": (C.c_int, [VP, C.POINTER(KpeTiling)]),
    "ltb_get_kpe_tiling": (C.c_int, [VP, C.c_int, C.POINTER(KpeTiling)]),
    "ltb_kpe_autotune": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(KpeTiling), PD]),
    "ltb_batch_create": (C.c_int, [VP, C.c_int, C.POINTER(VP), C.POINTER(VP)]),
    "ltb_batch_destroy": (None, [VP]),
    "ltb_batch_sizes": (C.c_int, [VP, PI, PI, PI]),
    "ltb_batch_stats": (C.c_int, [VP, C.POINTER(C.c_longlong), PD, PD]),
    "ltb_batch_set_tiling": (C.c_int, [VP, C.POINTER(KpeTiling)]),
    "ltb_batch_get_tiling": (C.c_int, [VP, C.POINTER(KpeTiling)]),
    "ltb_batch_autotune": (C.c_int, [VP, C.c_int, C.c_int, C.POINTER(KpeTiling), PD]),
    "ltb_batch_scc_run": (C.c_int, [VP, C.POINTER(SccParams), PD, PD, PI, PD, PD, PI,
                                    C.POINTER(SccResult)]),
    "ltb_mlp_num_params": (C.c_int, []),
    "ltb_mlp_create": (C.c_int, [VP, C.c_ulonglong, C.POINTER(VP)]),
    "ltb_mlp_destroy": (None, [VP]),
    "ltb_mlp_get_params": (C.c_int, [VP, C.POINTER(C.c_float)]),
    "ltb_mlp_set_params": (C.c_int, [VP, C.POINTER(C.c_float)]),
    "ltb_descriptors": (C.c_int, [VP, VP, C.POINTER(C.c_float)]),
    "ltb_mlp_set_data": (C.c_int, [VP, C.c_int, C.POINTER(C.c_float), PI, C.c_int, PI, PD, PD]),
    "ltb_mlp_fit_normalization": (C.c_int, [VP]),
    "ltb_mlp_forward": (C.c_int, [VP, PD]),
    "ltb_mlp_loss_grad": (C.c_int, [VP, PD, C.POINTER(C.c_float)]),
    "ltb_mlp_adam": (C.c_int, [VP, C.c_int, C.c_float, PD]),
    "ltb_mlp_predict": (C.c_int, [VP, VP, PD]),
}

MLP_NFEAT = 80
MLP_HIDDEN = 64
MLP_PER_SPECIES = MLP_HIDDEN * MLP_NFEAT + MLP_HIDDEN + 2 * (MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN) + MLP_HIDDEN + 1
MLP_TRAINABLE = 2 * MLP_PER_SPECIES


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python -c 'import "
                              "__graft_entry__ as g; g.build()' (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(PD)


def _vec(a, n, what):
    """a as a contiguous float64 vector of length n (the C side reads n doubles)"""
    v = np.ascontiguousarray(a, np.float64).reshape(-1)
    if v.size != n:
        raise ValueError(f"{what}: {v.size} values given, {n} expected")
    return v


def _i(a):
    return a.ctypes.data_as(PI)


def _sys_check(st):
    if st != LTB_OK:
        raise LtbError(st, lib().ltb_system_last_error().decode())


class System:
    """Host-side structure: geometry, neighbour list, H0 (BSR 4x4) and Gamma."""

    def __init__(self, handle):
        self._h = handle
        nat, nnzb, gnnz, ne = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _sys_check(lib().ltb_system_sizes(handle, C.byref(nat), C.byref(nnzb), C.byref(gnnz),
                                          C.byref(ne)))
        self.nat, self.nnzb, self.gnnz, self.nelectrons = nat.value, nnzb.value, gnnz.value, ne.value
        self.norb = 4 * self.nat

    @classmethod
    def quartz(cls, L, M, N, lattice=LATTICE_CONFIG, sigma_ang=0.0, seed=1):
        d = QuartzDesc(L, M, N, lattice, sigma_ang, seed)
        h = VP()
        _sys_check(lib().ltb_system_quartz(C.byref(d), C.byref(h)))
        return cls(h)

    @classmethod
    def from_coords(cls, Z, xyz, cell=None, periodic_mask=7):
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1)
        cp = None
        if cell is not None:
            cell = np.ascontiguousarray(cell, np.float64).reshape(9)
            cp = _d(cell)
        h = VP()
        _sys_check(lib().ltb_system_create(len(Z), _i(Z), _d(xyz), cp, periodic_mask, C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.ltb_system_destroy(self._h)
            self._h = None

    def coords(self):
        Z = np.zeros(self.nat, np.int32)
        xyz = np.zeros(3 * self.nat)
        _sys_check(lib().ltb_system_coords(self._h, _i(Z), _d(xyz)))
        return Z, xyz.reshape(-1, 3)

    def hamiltonian(self):
        rp = np.zeros(self.nat + 1, np.int32)
        col = np.zeros(self.nnzb, np.int32)
        val = np.zeros(self.nnzb * 16)
        _sys_check(lib().ltb_system_hamiltonian(self._h, _i(rp), _i(col), _d(val)))
        return rp, col, val.reshape(-1, 4, 4)

    def gamma(self):
        rp = np.zeros(self.nat + 1, np.int32)
        col = np.zeros(self.gnnz, np.int32)
        val = np.zeros(self.gnnz)
        _sys_check(lib().ltb_system_gamma(self._h, _i(rp), _i(col), _d(val)))
        return rp, col, val

    def neutral(self):
        p0 = np.zeros(self.nat)
        _sys_check(lib().ltb_system_neutral(self._h, _d(p0)))
        return p0

    def dense_h(self):
        rp, col, val = self.hamiltonian()
        H = np.zeros((self.norb, self.norb))
        for i in range(self.nat):
            for k in range(rp[i], rp[i + 1]):
                H[4 * i:4 * i + 4, 4 * col[k]:4 * col[k] + 4] = val[k]
        return H


def _io_check(st):
    if st != LTB_OK:
        raise LtbError(st, lib().ltb_io_last_error().decode())


def write_matlab_bsr(path, name, rowptr, col, val, bs, ncols=None):
    """Write a BSR (bs > 1) or CSR (bs = 1) matrix in the reference's MATLAB
    dump layout; ncols in block columns (default: square)."""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64).reshape(-1)
    n = len(rowptr) - 1
    _io_check(lib().ltb_write_matlab_bsr(str(path).encode(), name.encode(), n, n if ncols is None else ncols,
                                         bs, _i(rowptr), _i(col), _d(val)))


def write_matlab_vec(path, name, v):
    v = np.ascontiguousarray(v, np.float64)
    _io_check(lib().ltb_write_matlab_vec(str(path).encode(), name.encode(), len(v), _d(v)))


def read_matlab_mat(path):
    """(shape, rows, cols, vals) of a MATLAB spconvert dump, 0-based, file order."""
    nr, nc, nnz = C.c_longlong(), C.c_longlong(), C.c_longlong(0)
    _io_check(lib().ltb_read_matlab_mat(str(path).encode(), C.byref(nr), C.byref(nc), C.byref(nnz), None, None,
                                        None))
    r = np.zeros(nnz.value, np.int32)
    c = np.zeros(nnz.value, np.int32)
    v = np.zeros(nnz.value)
    _io_check(lib().ltb_read_matlab_mat(str(path).encode(), C.byref(nr), C.byref(nc), C.byref(nnz), _i(r), _i(c),
                                        _d(v)))
    return (nr.value, nc.value), r, c, v


def random_vectors(seed, nvec, n):
    """The seeded Gaussian vectors of the stochastic mode, [nvec, n]."""
    R = np.zeros((nvec, n))
    st = lib().ltb_random_vectors(seed, nvec, n, _d(R))
    if st:
        raise LtbError(st, "bad arguments")
    return R


def kpm_dos(moments, a, b, nsize, nelectrons, ntilde=None):
    """runKPM's DOS, mu and tightened bounds from trace moments (host)."""
    mom = np.ascontiguousarray(moments, np.float64)
    nt = 2 * len(mom) if ntilde is None else ntilde
    E, D = np.zeros(nt), np.zeros(nt)
    mu, lo, hi = C.c_double(), C.c_double(), C.c_double()
    st = lib().ltb_kpm_dos(_d(mom), len(mom), nt, a, b, nsize, nelectrons, _d(E), _d(D), C.byref(mu), C.byref(lo),
                           C.byref(hi))
    if st:
        raise LtbError(st, "bad arguments")
    return dict(energies=E, dos=D, mu=mu.value, emin=lo.value, emax=hi.value)


def write_dos(path, energies, dos):
    E = np.ascontiguousarray(energies, np.float64)
    D = np.ascontiguousarray(dos, np.float64)
    st = lib().ltb_write_dos(str(path).encode(), len(E), _d(E), _d(D))
    if st:
        raise LtbError(st, f"cannot write {path}")


def read_matlab_vec(path):
    n = C.c_longlong(0)
    _io_check(lib().ltb_read_matlab_vec(str(path).encode(), C.byref(n), None))
    v = np.zeros(n.value)
    _io_check(lib().ltb_read_matlab_vec(str(path).encode(), C.byref(n), _d(v)))
    return v


class Context:
    """Device context on one GPU (one HIP stream)."""

    def __init__(self, device=0):
        h = VP()
        st = lib().ltb_create(device, C.byref(h))
        if st != LTB_OK:
            raise LtbError(st, lib().ltb_last_error(None).decode())
        self._h = h
        self.system = None
        self._children = weakref.WeakSet()  # Batch / Mlp objects holding this context

    def close(self):
        """Destroy the context; its Batch and Mlp children are closed first
        (their device objects refer to the context's stream)."""
        if getattr(self, "_h", None) is not None:
            for child in list(self._children):
                child.close()
            lib().ltb_destroy(self._h)
            self._h = None

    def _handle(self):
        if getattr(self, "_h", None) is None:
            raise LtbError(LTB_ERR_STATE, "context is closed")
        return self._h

    def __del__(self):
        if _lib is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != LTB_OK:
            raise LtbError(st, lib().ltb_last_error(self._h).decode())

    def load(self, system):
        self._check(lib().ltb_load_system(self._h, system._h))
        self.system = system
        self.nat = system.nat
        self.norb = system.norb
        self.nnzb = system.nnzb
        self.nelectrons = system.nelectrons
        return self

    def load_coords(self, Z, xyz_bohr, cell9=None, periodic_mask=7):
        """Assemble neighbour list, H0 and Gamma on the device from coordinates."""
        Z = np.ascontiguousarray(Z, np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, np.float64).reshape(-1)
        cell = None if cell9 is None else _d(np.ascontiguousarray(cell9, np.float64).reshape(-1))
        sec = C.c_double()
        self._check(lib().ltb_load_coords(self._h, len(Z), _i(Z), _d(xyz), cell, periodic_mask,
                                          C.byref(sec)))
        nat, nnzb, gnnz, ne = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._check(lib().ltb_get_sizes(self._h, C.byref(nat), C.byref(nnzb), C.byref(gnnz), C.byref(ne)))
        self.system = None
        self.nat, self.norb, self.nnzb, self.nelectrons = nat.value, 4 * nat.value, nnzb.value, ne.value
        return sec.value

    def hamiltonian(self):
        """Host copy of the loaded H0: rowptr, col, val [nnzb, 4, 4]."""
        rp = np.zeros(self.nat + 1, np.int32)
        col = np.zeros(self.nnzb, np.int32)
        val = np.zeros(16 * self.nnzb)
        self._check(lib().ltb_get_hamiltonian(self._h, _i(rp), _i(col), _d(val)))
        return rp, col, val.reshape(-1, 4, 4)

    def gamma(self):
        nat, nnzb, gnnz, ne = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._check(lib().ltb_get_sizes(self._h, C.byref(nat), C.byref(nnzb), C.byref(gnnz), C.byref(ne)))
        rp = np.zeros(self.nat + 1, np.int32)
        col = np.zeros(gnnz.value, np.int32)
        val = np.zeros(gnnz.value)
        self._check(lib().ltb_get_gamma(self._h, _i(rp), _i(col), _d(val)))
        return rp, col, val

    def set_populations(self, p=None):
        """Coulomb shift from per-spin populations p; None keeps the
        populations the context holds (neutral after load, the last mixed
        ones after an SCC)."""
        pv = _vec(p, self.nat, "populations") if p is not None else None
        self._check(lib().ltb_set_populations(self._h, _d(pv) if pv is not None else None))

    def shift(self):
        v = np.zeros(self.nat)
        self._check(lib().ltb_get_shift(self._h, _d(v)))
        return v

    def set_neutral(self, p0):
        p0 = _vec(p0, self.nat, "neutral populations")
        self._check(lib().ltb_set_neutral(self._h, _d(p0)))

    def mixed_populations(self):
        """Per-spin populations the next SCC iteration starts from."""
        p = np.zeros(self.nat)
        self._check(lib().ltb_get_mixed_populations(self._h, _d(p)))
        return p

    def gershgorin(self):
        lo, hi = C.c_double(), C.c_double()
        self._check(lib().ltb_gershgorin(self._h, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def rescale(self, a, b):
        self._check(lib().ltb_rescale(self._h, a, b))

    def kpe_moments(self, M, precision=FP64, want_atoms=True):
        tot = np.zeros(M + 1)
        am = np.zeros((M + 1) * self.nat) if want_atoms else None
        self._check(lib().ltb_kpe_moments(self._h, M, precision, _d(am) if am is not None else None,
                                          _d(tot)))
        return (am.reshape(M + 1, self.nat) if am is not None else None), tot

    def density(self, M, kT_ev=0.05, K=8192, Ne=None):
        kT = kT_ev / HARTREE_EV if kT_ev > 0 else 0.0
        fp = FermiParams(M, kT, K, self.nelectrons if Ne is None else Ne)
        pop = np.zeros(self.nat)
        r = DensityResult()
        self._check(lib().ltb_density(self._h, C.byref(fp), _d(pop), C.byref(r)))
        return dict(pop=pop, mu=r.mu, mu_ha=r.mu_ha, a=r.a, b=r.b, nelec=r.nelec, eband=r.eband)

    def scc(self, M, guess=None, kT_ev=0.05, K=8192, mix=0.2, tol=1e-6, max_iter=200,
            precision=FP64, switch_tol=0.0, use_graph=0, check=True, verbose=0, M_lo=0, lo_tol=0.0):
        """ltb_scc_run.  M_lo > 0 selects the two-level SCC: most iterations
        evaluate the density at order M_lo corrected by the defect of the last
        full-order iteration; convergence is declared on full-order ones."""
        kT = kT_ev / HARTREE_EV if kT_ev > 0 else 0.0
        prm = SccParams(M, kT, K, mix, tol, max_iter, precision, switch_tol, use_graph, verbose, M_lo, lo_tol)
        pout = np.zeros(self.nat)
        hist = np.zeros(max_iter)
        r = SccResult()
        gv = _vec(guess, self.nat, "guess") if guess is not None else None
        st = lib().ltb_scc_run(self._h, C.byref(prm), _d(gv) if gv is not None else None, _d(pout), _d(hist),
                               C.byref(r))
        if check and st not in (LTB_OK,):
            self._check(st)
        return dict(status=st, pop=pout, q=2 * pout, iters=r.iterations, fp64_iters=r.fp64_iterations,
                    dq=r.dq, mu_ha=r.mu_ha, eband=r.eband, eband_fp=r.eband_fp,
                    graph_launches=r.graph_launches, seconds=r.seconds, lo_iters=r.lo_iterations,
                    hist=hist[:r.iterations])

    def kpe_weights_at(self, M, mu_ha, kT_ev=0.05, K=8192):
        kT = kT_ev / HARTREE_EV if kT_ev > 0 else 0.0
        w = np.zeros(M)
        self._check(lib().ltb_kpe_weights_at(self._h, M, mu_ha, kT, K, _d(w)))
        return w

    def block_moments(self, M, R, precision=FP64):
        """r^T T_m(H~) r for the rows r of R: ([nvec, M+1] moments, [M+1] trace)."""
        R = np.ascontiguousarray(R, np.float64)
        mom = np.zeros((R.shape[0], M + 1))
        tr = np.zeros(M + 1)
        self._check(lib().ltb_kpe_block_moments(self._h, M, R.shape[0], _d(R), precision, _d(mom), _d(tr)))
        return mom, tr

    def stochastic_moments(self, M, nvec, seed=1, precision=FP64):
        tr = np.zeros(M + 1)
        self._check(lib().ltb_kpe_stochastic(self._h, M, nvec, seed, precision, _d(tr)))
        return tr

    def density_stochastic(self, M, w, nvec, seed=1, precision=FP64):
        w = np.ascontiguousarray(w, np.float64)
        d = np.zeros(self.norb)
        self._check(lib().ltb_density_stochastic(self._h, M, _d(w), nvec, seed, precision, _d(d)))
        return d

    def trace_stochastic(self, M, w, nvec, seed=1, precision=FP64):
        """sum_r r^T P r / nvec over the seeded vectors (computeAlphas' trace)"""
        w = np.ascontiguousarray(w, np.float64)
        t = C.c_double()
        self._check(lib().ltb_trace_stochastic(self._h, M, _d(w), nvec, seed, precision, C.byref(t)))
        return t.value

    def keep_orbital_moments(self, on=True):
        self._check(lib().ltb_set_kpe_orbital_moments(self._h, 1 if on else 0))

    def set_gamma_onsite(self, gamma_aa):
        """Gamma_AA per atom (ltb_set_gamma_onsite); hubbard_u(Z) gives the +U values."""
        self._check(lib().ltb_set_gamma_onsite(self._h, _d(_vec(gamma_aa, self.nat, "gamma_aa"))))

    def set_light_cone(self, on=True):
        """Light-cone restriction of the unit-vector recursion (default on)."""
        self._check(lib().ltb_set_kpe_light_cone(self._h, 1 if on else 0))

    def kpe_weights(self, M):
        w = np.zeros(M)
        self._check(lib().ltb_kpe_weights(self._h, M, _d(w)))
        return w

    def orbital_moments(self, M, j0, j1):
        """[T_m(H~)]_jj for m = 0..M and j0 <= j < j1: [M+1, j1-j0]"""
        out = np.zeros((M + 1) * (j1 - j0))
        self._check(lib().ltb_get_orbital_moments(self._h, M, j0, j1, _d(out)))
        return out.reshape(M + 1, j1 - j0)

    def density_diag(self):
        d = np.zeros(self.norb)
        self._check(lib().ltb_density_diag(self._h, _d(d)))
        return d

    def kpe_apply(self, M, w, X, precision=FP64):
        X = np.ascontiguousarray(X, np.float64)
        if X.ndim == 1:
            X = X[None, :]
        w = np.ascontiguousarray(w, np.float64)
        Y = np.zeros_like(X)
        self._check(lib().ltb_kpe_apply(self._h, M, _d(w), X.shape[0], _d(X), _d(Y), precision))
        return Y

    def sp2_dense(self, Ne=None, max_iter=100, fp64_tail=2):
        diag = np.zeros(self.norb)
        pop = np.zeros(self.nat)
        r = Sp2Result()
        self._check(lib().ltb_sp2_dense(self._h, self.nelectrons if Ne is None else Ne,
                                        max_iter, fp64_tail, _d(diag), _d(pop), C.byref(r)))
        return dict(diag=diag, pop=pop, iters=r.iterations, fp64_iters=r.fp64_iterations,
                    trace=r.trace, idem=r.idem, gemm_seconds=r.gemm_seconds, gemm_flops=r.gemm_flops)

    def kpe_stats(self):
        n, s, b = C.c_int(), C.c_double(), C.c_double()
        self._check(lib().ltb_kpe_stats(self._h, C.byref(n), C.byref(s), C.byref(b)))
        return dict(launches=n.value, seconds=s.value, bytes=b.value)

    def set_timing(self, on=True):
        self._check(lib().ltb_set_kpe_timing(self._h, 1 if on else 0))

    def set_tiling(self, chunk_cols=0, cols_per_lane=0, schedule=SCHED_AUTO, blocks_per_cu=0,
                   atoms_per_wave=0, nontemporal=-1, order=0, lo_chunk_cols=0):
        t = KpeTiling(chunk_cols, cols_per_lane, schedule, blocks_per_cu, atoms_per_wave, nontemporal, order,
                      lo_chunk_cols)
        self._check(lib().ltb_set_kpe_tiling(self._h, C.byref(t)))

    def get_tiling(self, precision=FP32_ACC64):
        t = KpeTiling()
        self._check(lib().ltb_get_kpe_tiling(self._h, precision, C.byref(t)))
        return t.as_dict()

    def autotune(self, precision=FP32_ACC64, trial_steps=3):
        t = KpeTiling()
        g = C.c_double()
        self._check(lib().ltb_kpe_autotune(self._h, precision, trial_steps, C.byref(t), C.byref(g)))
        return t.as_dict(), g.value

    def descriptors(self, system):
        f = np.zeros((system.nat, MLP_NFEAT), np.float32)
        self._check(lib().ltb_descriptors(self._h, system._h, f.ctypes.data_as(C.POINTER(C.c_float))))
        return f


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Batch:
    """Independent structures of uneven size solved together on one context:
    each keeps its own bounds, mu, mixing, precision switch and convergence,
    while every Chebyshev step of all of them is one launch."""

    def __init__(self, ctx, systems):
        self.ctx = ctx
        arr = (VP * len(systems))(*[s._h for s in systems])
        h = VP()
        ctx._check(lib().ltb_batch_create(ctx._handle(), len(systems), arr, C.byref(h)))
        self._h = h
        ctx._children.add(self)
        S, nat = C.c_int(), C.c_int()
        off = np.zeros(len(systems) + 1, np.int32)
        ctx._check(lib().ltb_batch_sizes(h, C.byref(S), C.byref(nat), _i(off)))
        self.S, self.nat, self.offsets = S.value, nat.value, off

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ltb_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def scc(self, M, guess=None, kT_ev=0.05, K=8192, mix=0.2, tol=1e-6, max_iter=200,
            precision=FP64, switch_tol=0.0, use_graph=0, check=True, verbose=0, M_lo=0, lo_tol=0.0):
        """ltb_batch_scc_run.  M_lo > 0: structures recursing in FP32 run the
        two-level SCC until a full-order iteration has |dq| < switch_tol, then
        plain FP64 iterations."""
        kT = kT_ev / HARTREE_EV if kT_ev > 0 else 0.0
        prm = SccParams(M, kT, K, mix, tol, max_iter, precision, switch_tol, use_graph, verbose, M_lo, lo_tol)
        pout = np.zeros(self.nat)
        its = np.zeros(self.S, np.int32)
        ebfp = np.zeros(self.S)
        mu = np.zeros(self.S)
        status = np.zeros(self.S, np.int32)
        r = SccResult()
        gv = _vec(guess, self.nat, "guess") if guess is not None else None
        st = lib().ltb_batch_scc_run(self._batch(), C.byref(prm), _d(gv) if gv is not None else None, _d(pout),
                                     _i(its), _d(ebfp), _d(mu), _i(status), C.byref(r))
        if check:
            self.ctx._check(st)
        return dict(status=st, pop=pout, q=2 * pout, iters=its, eband_fp=ebfp, mu_ha=mu,
                    struct_status=status, batch_iters=r.iterations, fp64_iters=r.fp64_iterations,
                    graph_launches=r.graph_launches, seconds=r.seconds, lo_iters=r.lo_iterations)

    def tiling(self):
        t = KpeTiling()
        self.ctx._check(lib().ltb_batch_get_tiling(self._batch(), C.byref(t)))
        return {k: getattr(t, k) for k, _ in KpeTiling._fields_}

    def set_tiling(self, atoms_per_wave=0, blocks_per_cu=0, nontemporal=-1):
        t = KpeTiling(0, 0, 0, blocks_per_cu, atoms_per_wave, nontemporal, 0)
        self.ctx._check(lib().ltb_batch_set_tiling(self._batch(), C.byref(t)))

    def autotune(self, precision=FP32_ACC64, trial_steps=3):
        t, g = KpeTiling(), C.c_double()
        self.ctx._check(lib().ltb_batch_autotune(self._batch(), precision, trial_steps, C.byref(t), C.byref(g)))
        return {k: getattr(t, k) for k, _ in KpeTiling._fields_}, g.value

    def stats(self):
        n, sec, by = C.c_longlong(), C.c_double(), C.c_double()
        self.ctx._check(lib().ltb_batch_stats(self._batch(), C.byref(n), C.byref(sec), C.byref(by)))
        return dict(launches=n.value, seconds=sec.value, bytes=by.value)

    def _batch(self):
        if getattr(self, "_h", None) is None:
            raise LtbError(LTB_ERR_STATE, "batch is closed")
        return self._h

    def split(self, x):
        """Per-structure views of a concatenated per-atom array."""
        return [x[self.offsets[k]:self.offsets[k + 1]] for k in range(self.S)]


class Mlp:
    """Per-species 80-64-64-64-1 SiLU network on the device (FP32) with Adam."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = VP()
        ctx._check(lib().ltb_mlp_create(ctx._handle(), seed, C.byref(h)))
        self._h = h
        ctx._children.add(self)
        self.nparams = lib().ltb_mlp_num_params()

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().ltb_mlp_destroy(self._h)
            self._h = None

    def _mlp(self):
        if getattr(self, "_h", None) is None:
            raise LtbError(LTB_ERR_STATE, "network is closed")
        return self._h

    def __del__(self):
        if _lib is not None:
            self.close()

    def params(self):
        p = np.zeros(self.nparams, np.float32)
        self.ctx._check(lib().ltb_mlp_get_params(self._mlp(), _f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self.ctx._check(lib().ltb_mlp_set_params(self._mlp(), _f(p)))

    def set_data(self, feat, Z, seg, qtot, target=None):
        Z = np.ascontiguousarray(Z, np.int32).reshape(-1)
        n = len(Z)
        feat = np.ascontiguousarray(feat, np.float32)
        if feat.shape != (n, MLP_NFEAT):
            raise ValueError(f"features: shape {feat.shape}, ({n}, {MLP_NFEAT}) expected")
        seg = np.ascontiguousarray(seg, np.int32).reshape(-1)
        if len(seg) < 2:
            raise ValueError("segments: at least one structure [seg[0], seg[1]) expected")
        qtot = _vec(qtot, len(seg) - 1, "electrons per structure")
        self._target = None if target is None else _vec(target, n, "target charges")
        self.n = len(Z)
        self.ctx._check(lib().ltb_mlp_set_data(self._mlp(), len(Z), _f(feat), _i(Z), len(seg) - 1, _i(seg),
                                               _d(qtot), _d(self._target) if self._target is not None else None))

    def fit_normalization(self):
        self.ctx._check(lib().ltb_mlp_fit_normalization(self._mlp()))

    def forward(self):
        q = np.zeros(self.n)
        self.ctx._check(lib().ltb_mlp_forward(self._mlp(), _d(q)))
        return q

    def loss_grad(self):
        loss = C.c_double()
        g = np.zeros(MLP_TRAINABLE, np.float32)
        self.ctx._check(lib().ltb_mlp_loss_grad(self._mlp(), C.byref(loss), _f(g)))
        return loss.value, g

    def adam(self, steps, lr=1e-3):
        hist = np.zeros(steps)
        self.ctx._check(lib().ltb_mlp_adam(self._mlp(), steps, lr, _d(hist)))
        return hist

    def predict(self, system):
        p = np.zeros(system.nat)
        self.ctx._check(lib().ltb_mlp_predict(self._mlp(), system._h, _d(p)))
        return p
