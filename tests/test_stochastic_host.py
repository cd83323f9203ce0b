"""Host parts of the stochastic-trace mode: the seeded Gaussian vectors and
runKPM's DOS / mu / bound tightening (ltb_kpm_dos), against the oracle's
literal restatement and against what the spectrum fixes."""
import re

import numpy as np
import pytest

import largetb_amd as L
import oracle as O


def test_gaussian_stream_is_the_displacement_generator():
    """The vectors come from the stream that displaces the atoms: the
    displacement of a quartz cell divided by sigma reproduces it."""
    sig, seed = 0.1, 5
    d = (O.System.quartz(2, 2, 2, sigma_ang=sig, seed=seed).xyz - O.System.quartz(2, 2, 2).xyz).reshape(-1)
    g = O.gaussian_stream(seed, d.size)
    assert np.abs(d / (sig * 1.88973) - g).max() < 1e-12
    R = L.random_vectors(seed, 3, 72)
    assert np.array_equal(R.reshape(-1), O.gaussian_stream(seed, 216))


@pytest.fixture(scope="module")
def c0_exact():
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    lo, hi = o.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    nm = 400
    mom = o.diag_moments(a, b, nm).sum(0)  # exact Tr T_m, m < nm
    return o, a, b, mom


def test_dos_matches_literal_restatement(c0_exact):
    o, a, b, mom = c0_exact
    ne = 384.0
    got = L.kpm_dos(mom, a, b, o.norb, ne)
    ref = O.kpm_dos(mom, a, b, o.norb, ne)
    assert np.abs(got["dos"] - ref["dos"]).max() < 1e-9 * np.abs(ref["dos"]).max()
    assert np.array_equal(got["energies"], ref["energies"])
    assert got["mu"] == ref["mu"] and got["emin"] == ref["emin"] and got["emax"] == ref["emax"]


def test_dos_against_the_spectrum(c0_exact):
    """The DOS integrates to the orbital count, mu from the DOS lies in the
    gap, and the tightened bounds sit near the extreme eigenvalues."""
    o, a, b, mom = c0_exact
    r = L.kpm_dos(mom, a, b, o.norb, 384.0)
    x = (r["energies"] - b) / a
    area = np.sum((x[:-1] - x[1:]) * (r["dos"][1:] + r["dos"][:-1]) / 2)
    assert area == pytest.approx(o.norb, rel=2e-3)
    e = np.linalg.eigvalsh(o.dense_h())
    homo, lumo = e[191], e[192]
    assert homo - 0.02 < r["mu"] < lumo + 0.02
    assert abs(r["emin"] - e[0]) < 0.05 * a and abs(r["emax"] - e[-1]) < 0.05 * a


def test_dos_file_layout(tmp_path, c0_exact):
    o, a, b, mom = c0_exact
    r = L.kpm_dos(mom, a, b, o.norb, 384.0)
    L.write_dos(tmp_path / "dosfile.txt", r["energies"], r["dos"])
    lines = (tmp_path / "dosfile.txt").read_text().split("\n")[:-1]
    assert len(lines) == 2 * len(mom)
    assert all(re.fullmatch(r"-?\d+\.\d{6}, -?\d+\.\d{6}  ", l) for l in lines)
