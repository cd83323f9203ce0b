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


# ---- the replay of the reference's estimator from exact eigenvalues ----------------


def test_replay_eigen_expansion_equals_the_estimator():
    """The replay's premise, on the 72-atom reference-lattice cell: for
    Gaussian vectors r, sum_r r^T P r with P applied by the oracle's getKPE op
    sequence equals sum_k w_k f(e_k), w_k = sum_r (v_k^T r)^2, where f is the
    replay's occupation polynomial at the same mu."""
    from oracle import stochastic_replay as SR
    o = O.System.quartz(2, 2, 2, mode=O.MODE_REFERENCE)
    e, V = np.linalg.eigh(o.dense_h())
    lo, hi = o.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    muval = ((e[191] + e[192]) / 2 - b) / a  # mid-gap, 384 electrons
    w = O.kpe_weights(O.step_coeffs(200, muval), O.gmM(200))
    f = SR.step_occupation((e - b) / a, muval)
    R = L.random_vectors(7, 5, o.norb)
    est = sum(r @ o.kpe_apply(a, b, w, r) for r in R)
    C = V.T @ R.T
    assert abs(est - ((C ** 2).sum(1) * f).sum()) < 1e-9 * abs(est)


def test_replay_bound_narrowing_keeps_the_gershgorin_frame():
    """runKPM's DOS is summed over the vectors, and 1/sqrt(1 - x^2) lifts the
    outermost of the 2000 nodes above 0.001 per orbital for 500 vectors: the
    narrowed bounds (:1695-1712) are the outermost nodes of the Gershgorin
    frame.  For one vector the same rule does narrow the frame."""
    from oracle import stochastic_replay as SR
    o = O.System.quartz(3, 3, 3, mode=O.MODE_REFERENCE)
    e = np.linalg.eigvalsh(o.dense_h())
    lo, hi = o.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    edge = np.cos(np.pi * 0.5 / 2000)
    for r in SR.replay(e, (lo, hi), 48.0 * 27, ndraw=2):
        assert r["emin"] == pytest.approx(b - edge * a, abs=1e-12)
        assert r["emax"] == pytest.approx(b + edge * a, abs=1e-12)
        assert lo < e[0] and e[-1] < hi
    one = SR.replay(e, (lo, hi), 48.0 * 27, nrandoms=1, ndraw=1)[0]
    assert one["emin"] > b - 0.99 * a and one["emax"] < b + 0.99 * a
