"""MATLAB ASCII dumps (the reference's PETSc viewer format), host only.

The reference-lattice 4x4x4 H written by ltb_write_matlab_bsr lists exactly
the entries of the reference's own H.m (same 32,768 (row, col) pairs in the
same order, explicit zeros of the 4x4 blocks included) with values within
2e-15; Gamma likewise against GAMMA_AB.m.  Values round-trip bit for bit."""
import numpy as np
import pytest

import largetb_amd as L
from tests.golden.make_reference_fixtures import read_petsc_mat, read_petsc_vec


def test_h_dump_reproduces_reference_file(tmp_path, ref_dumps):
    s = L.System.quartz(4, 4, 4, lattice=L.LATTICE_REFERENCE)
    rp, col, val = s.hamiltonian()
    path = tmp_path / "H.m"
    L.write_matlab_bsr(path, "H", rp, col, val, 4)
    shape, r, c, v = L.read_matlab_mat(path)
    assert shape == tuple(ref_dumps["H_shape"])
    assert np.array_equal(r, ref_dumps["H_row"]) and np.array_equal(c, ref_dumps["H_col"])
    assert np.abs(v - ref_dumps["H_val"]).max() < 2e-15
    # the fixtures' own parser (written for the reference's files) reads it the same way
    size, r2, c2, v2 = read_petsc_mat(str(path))
    assert size == shape and np.array_equal(r2, r) and np.array_equal(c2, c) and np.array_equal(v2, v)


def test_gamma_dump_reproduces_reference_file(tmp_path, ref_dumps):
    s = L.System.quartz(4, 4, 4, lattice=L.LATTICE_REFERENCE)
    rp, col, val = s.gamma()
    path = tmp_path / "GAMMA_AB.m"
    L.write_matlab_bsr(path, "GAMMA_AB", rp, col, val, 1)
    shape, r, c, v = L.read_matlab_mat(path)
    assert shape == tuple(ref_dumps["GAMMA_AB_shape"])
    assert np.array_equal(r, ref_dumps["GAMMA_AB_row"]) and np.array_equal(c, ref_dumps["GAMMA_AB_col"])
    assert np.abs(v - ref_dumps["GAMMA_AB_val"]).max() < 1e-15


def test_round_trip_is_exact(tmp_path):
    s = L.System.quartz(2, 3, 2, sigma_ang=0.1, seed=9)
    rp, col, val = s.hamiltonian()
    L.write_matlab_bsr(tmp_path / "h.m", "Hx", rp, col, val, 4)
    shape, r, c, v = L.read_matlab_mat(tmp_path / "h.m")
    dense = np.zeros(shape)
    dense[r, c] = v
    assert np.array_equal(dense, s.dense_h())
    x = np.random.default_rng(1).normal(size=77) * 10.0 ** np.arange(-38, 39)
    L.write_matlab_vec(tmp_path / "q.m", "Vec_netCharge", x)
    assert np.array_equal(L.read_matlab_vec(tmp_path / "q.m"), x)
    assert np.array_equal(read_petsc_vec(str(tmp_path / "q.m")), x)


def test_layout_lines(tmp_path):
    L.write_matlab_bsr(tmp_path / "g.m", "G", [0, 1, 2], [0, 1], [-1.0, 0.5], 1)
    lines = (tmp_path / "g.m").read_text().splitlines()
    assert lines[:6] == ["%Mat Object: 1 MPI process", "%  type: seqaij", "% Size = 2 2 ", "% Nonzeros = 2 ",
                         "zzz = zeros(2,3);", "zzz = ["]
    assert lines[6] == "1 1  -1.0000000000000000e+00"
    assert lines[-2:] == ["];", " Mat_G = spconvert(zzz);"]


def test_errors(tmp_path):
    with pytest.raises(L.LtbError):
        L.write_matlab_vec(tmp_path / "v.m", "1bad name", np.zeros(3))
    with pytest.raises(L.LtbError):
        L.read_matlab_mat(tmp_path / "missing.m")
    L.write_matlab_vec(tmp_path / "v.m", "v", np.ones(3))
    with pytest.raises(L.LtbError):
        L.read_matlab_mat(tmp_path / "v.m")  # a vector file is not a matrix
