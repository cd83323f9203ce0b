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
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ref = o.scc(200, tol=1e-6, maxit=300)
    assert abs(out["iterations"] - ref["iters"]) <= 1
    assert np.abs(q - ref["q"]).max() < 1e-8
    assert out["total_charge"] == pytest.approx(384.0, abs=1e-8)


@pytest.mark.gpu
def test_sp2_depth_one_through_getSP2(tmp_path):
    """maxsp2 = 2: diag(P1) with P1 = P0^2 or 2 P0 - P0^2 chosen by the trace
    rule of computeAlphas, P0 the KPE density -- applied to every unit vector
    by the recursive getSP2, as the reference's exact mode does."""
    r, out, q = run(["2", "2", "2", "--sigma", "0.05", "--seed", "1", "--moments", "200",
                     "--maxsp2", "2", "--max-scf", "1"], tmp_path, timeout=600)
    assert r.returncode in (0, 5), r.stderr
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    d = o.density(200)
    H = o.dense_h()
    e, V = np.linalg.eigh(H)
    x = (e - d["b"]) / d["a"]
    c = O.fd_coeffs(200, d["mu"], 0.05 / O.HARTREE_EV / d["a"], 8192)
    f = np.polynomial.chebyshev.chebval(x, O.kpe_weights(c, O.gmM(200)))
    alpha = -1.0 if 2 * f.sum() > o.nelectrons else 1.0
    f1 = (1 + alpha) * f - alpha * f * f
    pop = ((V ** 2) @ f1).reshape(-1, 4).sum(1)
    assert np.abs(q - 2 * pop).max() < 1e-8
