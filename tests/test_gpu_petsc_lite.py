"""The reference-named entry points (include/ltb_petsc.hpp) driven by ltb_scf,
the counterpart of the reference's `./scf L M N` (withSP2_deltaH.cpp main)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT

DRIVER = os.path.join(ROOT, "largetb_amd", "ltb_scf")


def run(args, tmp_path, timeout=240):
    qfile = tmp_path / "q.txt"
    env = dict(os.environ, LTB_SCF_CHARGES=str(qfile))
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=timeout, env=env)
    last = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert last, r.stdout + r.stderr
    return r, json.loads(last[-1]), np.loadtxt(qfile)


def test_driver_usage_without_gpu():
    r = subprocess.run([DRIVER], capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_scf_driver_matches_oracle(tmp_path):
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "200",
                     "--max-scf", "300"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "scfiter : 0 , the value of the scfDelta is" in r.stdout
    # the reference's phase timers, and one runKPM + density per iteration
    assert "time KPM : " in r.stdout and "time SP2 : " in r.stdout
    assert out["runKPM"]["calls"] == out["iterations"] == out["computeDensityMatrix"]["calls"]
    assert out["runKPM"]["seconds"] > 0
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ref = o.scc(200, tol=1e-6, maxit=300)
    assert abs(out["iterations"] - ref["iters"]) <= 1
    assert np.abs(q - ref["q"]).max() < 1e-8
    assert out["total_charge"] == pytest.approx(384.0, abs=1e-8)


@pytest.mark.gpu
def test_scf_driver_two_level(tmp_path):
    """ltb_scf --two-level 32: the same fixed point as the oracle's plain SCC
    (the two-level SCC converges only on a full-order iteration), with the
    reference's per-iteration lines, in FP64 and with --fp32."""
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ref = o.scc(200, tol=1e-10, maxit=600)
    for extra in ([], ["--fp32"]):
        r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "200", "--max-scf", "1000",
                         "--two-level", "32"] + extra, tmp_path)
        assert r.returncode == 0, r.stderr
        assert out["status"] == 0 and out["iterations"] > 100
        assert r.stdout.count("scfiter : ") == out["iterations"]
        assert np.abs(q - ref["q"]).max() < 1e-6
        assert out["total_charge"] == pytest.approx(384.0, abs=1e-8)
    r = subprocess.run([DRIVER, "2", "2", "2", "--moments", "200", "--two-level", "200"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 1 and "two-level" in r.stderr


def p0_eigen(M=200):
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    d = o.density(M)
    e, V = np.linalg.eigh(o.dense_h())
    c = O.fd_coeffs(M, d["mu"], 0.05 / O.HARTREE_EV / d["a"], 8192)
    f = np.polynomial.chebyshev.chebval((e - d["b"]) / d["a"], O.kpe_weights(c, O.gmM(M)))
    return o, V, f


def sp2_levels(f, alphas):
    """getSP2's recursion on eigenvalues: level k maps x -> (1 + a_k) x - a_k x^2"""
    for a in alphas:
        f = (1 + a) * f - a * f * f
    return f


@pytest.mark.gpu
def test_sp2_depth_four_through_getSP2(tmp_path):
    """maxsp2 = 5, withSP2.cpp's depth (SP2:1110), with the level pattern
    +1, -1, +1, -1: sixteen getKPE calls per unit vector through the recursive
    getSP2 (exact mode), against the SP2 polynomial of the KPE density's
    eigenvalues."""
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "100",
                     "--maxsp2", "5", "--max-scf", "1", "--sp2-alphas", "1,-1,1,-1"], tmp_path, timeout=600)
    assert r.returncode in (0, 5), r.stderr
    o, V, f = p0_eigen(100)
    pop = ((V ** 2) @ sp2_levels(f, [1, -1, 1, -1])).reshape(-1, 4).sum(1)
    assert np.abs(q - 2 * pop).max() < 1e-8


@pytest.mark.gpu
def test_sp2_with_the_doubled_kpe_output(tmp_path):
    """withSP2.cpp's getKPE returns 2 P x (VecScale(outVector, 2.0), SP2:939):
    --kpe-times-two at depth 2 (levels -1, +1) gives the SP2 polynomial of
    2 P0."""
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "100",
                     "--maxsp2", "3", "--max-scf", "1", "--sp2-alphas", "-1,1", "--kpe-times-two"], tmp_path,
                    timeout=600)
    assert r.returncode in (0, 5), r.stderr
    o, V, f = p0_eigen(100)
    pop = ((V ** 2) @ sp2_levels(2 * f, [-1, 1])).reshape(-1, 4).sum(1)
    assert np.abs(q - 2 * pop).max() < 1e-8 * max(1.0, np.abs(pop).max())


@pytest.mark.gpu
def test_sp2_depth_one_through_getSP2(tmp_path):
    """maxsp2 = 2 with alpha = -1: diag(P0^2), P0 the KPE density, applied to
    every unit vector through the recursive getSP2 (the reference's exact
    mode), against eigenpairs."""
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "200",
                     "--maxsp2", "2", "--max-scf", "1", "--sp2-alpha", "-1"], tmp_path, timeout=600)
    assert r.returncode in (0, 5), r.stderr
    o, V, f = p0_eigen()
    pop = ((V ** 2) @ (f * f)).reshape(-1, 4).sum(1)
    assert np.abs(q - 2 * pop).max() < 1e-8


@pytest.mark.gpu
def test_sp2_rule_stops_when_trace_is_exact(tmp_path):
    """computeAlphas' rule (:1278): with mu bisected on exact moments,
    |2 tr P0 - Ne| < 1e-3, so alpha stays 0 and depth 1 returns P0."""
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "200",
                     "--maxsp2", "2", "--max-scf", "1"], tmp_path, timeout=600)
    assert r.returncode in (0, 5), r.stderr
    assert "sp2count is 0, and trace is 192.0" in r.stdout
    o, V, f = p0_eigen()
    pop = ((V ** 2) @ f).reshape(-1, 4).sum(1)
    assert np.abs(q - 2 * pop).max() < 1e-8


@pytest.mark.gpu
def test_driver_writes_the_reference_dumps(tmp_path, ref_dumps):
    """--dump writes the reference's MATLAB files at the reference's points:
    H.m and GAMMA_AB.m of the reference-lattice 4x4x4 cell equal the
    reference's own dumps entry for entry; deltaH.m / netCharge.m /
    GammaQ_vector.m of the last SCC iteration satisfy GammaQ = Gamma . netCharge
    and deltaH = diag(repeat(GammaQ, 4))."""
    import largetb_amd as L
    d = tmp_path / "dump"
    d.mkdir()
    r, out, q = run(["4", "4", "4", "--reference", "--max-scf", "2", "--dump", str(d)], tmp_path)
    assert r.returncode in (0, 5), r.stderr
    shape, rr, cc, vv = L.read_matlab_mat(d / "H.m")
    assert np.array_equal(rr, ref_dumps["H_row"]) and np.array_equal(cc, ref_dumps["H_col"])
    assert np.abs(vv - ref_dumps["H_val"]).max() < 2e-15
    shape, rr, cc, vv = L.read_matlab_mat(d / "GAMMA_AB.m")
    assert np.array_equal(rr, ref_dumps["GAMMA_AB_row"]) and np.array_equal(cc, ref_dumps["GAMMA_AB_col"])
    G = np.zeros(shape)
    G[rr, cc] = vv
    dq = L.read_matlab_vec(d / "netCharge.m")
    gq = L.read_matlab_vec(d / "GammaQ_vector.m")
    assert np.abs(dq).max() > 1e-3  # second iteration: charges have moved
    assert np.abs(G @ dq - gq).max() < 1e-12
    shape, rr, cc, vv = L.read_matlab_mat(d / "deltaH.m")
    assert shape == (4 * len(gq), 4 * len(gq)) and len(vv) == 16 * len(gq)
    D = np.zeros(shape)
    D[rr, cc] = vv
    assert np.array_equal(D, np.diag(np.repeat(gq, 4)))


@pytest.mark.gpu
def test_driver_stochastic_mode(tmp_path):
    """`ltb_scf L M N nRandoms`: runKPM's stochastic DOS and mu, and the
    random-vector diagonal.  On the neutral first iteration mu (the DOS-area
    rule puts it at the valence-band edge) lands at the gap of H; the charges
    scatter, by the estimator's noise, around the exact populations at that
    same mu; dosfile.txt holds the 2 x 1000-point DOS."""
    d = tmp_path / "dump"
    d.mkdir()
    r, out, q = run(["2", "2", "2", "64", "--sigma", "0.05", "--seed", "1", "--max-scf", "1", "--dump", str(d)],
                    tmp_path)
    assert r.returncode in (0, 5), r.stderr
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    e = np.linalg.eigvalsh(o.dense_h())
    assert e[191] - 0.02 < out["mu_ha"] < e[192] + 0.02
    # exact populations at ltb_scf's mu: per-atom moments times the weights
    lo, hi = o.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    c = O.fd_coeffs(200, (out["mu_ha"] - b) / a, 0.05 / O.HARTREE_EV / a, 8192)
    w = O.kpe_weights(c, O.gmM(200))
    exact = 2 * (o.diag_moments(a, b, 200) @ w).reshape(-1, 4).sum(1)
    assert np.abs(q - exact).mean() < 0.3 and np.abs(q - exact).max() < 2.0
    assert abs(q.sum() - exact.sum()) < 12.0
    assert len((d / "dosfile.txt").read_text().splitlines()) == 2000


@pytest.mark.gpu
def test_driver_reference_literal_mode(tmp_path):
    """--reference-literal with mixRatio 0 (the reference's default, :56):
    H_total never changes, so the loop stops on its first iteration with a
    zero change between successive outputs, the charges are the populations
    of H itself, and netCharge.m holds them minus the literal neutral charges
    (Si 2, O 0 per spin, :714-720)."""
    import largetb_amd as L
    d = tmp_path / "dump"
    d.mkdir()
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--reference-literal", "--mix", "0",
                     "--tol", "0.1", "--max-scf", "20", "--dump", str(d)], tmp_path)
    assert r.returncode == 0, r.stderr
    assert out["iterations"] == 1 and out["delta"] < 1e-12
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ref = o.density(200)["pop"]
    assert np.abs(q - 2 * ref).max() < 1e-8
    net = L.read_matlab_vec(d / "netCharge.m")
    assert np.abs(net - (ref - np.where(o.Z == 14, 2.0, 0.0))).max() < 1e-8
