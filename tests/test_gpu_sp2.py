"""Dense SP2 on MFMA (FP32 bulk + FP64 tail) against the eigen-projector, the
oracle's FP64 dense SP2 and the KPE density matrix."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


def test_sp2_c0_matches_projector_and_oracle(ctx):
    s = L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    H = o.dense_h()
    e, V = np.linalg.eigh(H)
    nocc = int(o.nelectrons // 2)
    P = V[:, :nocc] @ V[:, :nocc].T
    lo, hi = o.gershgorin()
    X, it, _ = O.sp2_dense((hi * np.eye(o.norb) - H) / (hi - lo), nocc, maxit=100, idem_tol=1e-10)
    # FP64 throughout: the projector to rounding, same iteration count as the oracle
    r64 = ctx.sp2_dense(max_iter=0, fp64_tail=100)
    assert np.abs(r64["diag"] - np.diag(P)).max() < 1e-9
    assert np.abs(r64["diag"] - np.diag(X)).max() < 1e-9
    assert np.abs(r64["pop"] - np.diag(P).reshape(-1, 4).sum(1)).max() < 4e-9
    assert abs(r64["iters"] - it) <= 2
    # FP32 bulk + 2 FP64 iterations: idempotent (integer trace), subspace at FP32 rounding
    r = ctx.sp2_dense(max_iter=100, fp64_tail=2)
    assert r["trace"] == pytest.approx(nocc, abs=1e-8)
    assert r["fp64_iters"] == 2 and r["iters"] < 60
    assert np.abs(r["diag"] - np.diag(P)).max() < 5e-5


def test_sp2_with_coulomb_shift(ctx):
    s = L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ctx.load(s)
    rng = np.random.default_rng(2)
    p = o.p0 + 0.05 * rng.standard_normal(o.nat)
    ctx.set_populations(p)
    r = ctx.sp2_dense(max_iter=0, fp64_tail=100)
    v = o.dense_gamma() @ (p - o.p0)
    H = o.dense_h(np.repeat(v, 4))
    e, V = np.linalg.eigh(H)
    nocc = int(o.nelectrons // 2)
    assert np.abs(r["diag"] - (V[:, :nocc] ** 2).sum(1)).max() < 1e-9


def test_sp2_c1_against_kpe(ctx):
    """18,432 orbitals: SP2 (T = 0 projector) vs KPE M=512 at kT = 0.05 eV
    (gap 6.6 eV >> kT): populations agree to the KPE truncation (<2e-3 e)."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    r = ctx.sp2_dense(max_iter=100, fp64_tail=2)
    assert r["trace"] == pytest.approx(s.nelectrons / 2, abs=1e-6)
    assert r["idem"] < 1e-6
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    ctx.kpe_moments(512, L.FP64)
    d = ctx.density(512)
    assert np.abs(r["pop"] - d["pop"]).max() < 2e-3
    tflops = r["gemm_flops"] / r["gemm_seconds"] / 1e12
    print(f"SP2 18432: {r['iters']} iterations, {tflops:.1f} TFLOP/s over the GEMMs")
