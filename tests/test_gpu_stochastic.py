"""Stochastic-trace mode on the device (ltb_kpe_block_moments,
ltb_kpe_stochastic, ltb_density_stochastic) against dense FP64 references.

With unit vectors the block recursion is the exact trace (pins the
explicit-start path); with the seeded Gaussian vectors it equals
sum_r r^T T_m r computed densely (1e-10 relative in FP64); the Gaussian
estimate of Tr T_m stays within 5 standard deviations (Var r^T A r = 2 |A|_F^2)."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0(ctx):
    s = L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    ctx.rescale(a, b)
    Ht = (o.dense_h() - b * np.eye(o.norb)) / a
    return s, o, a, b, Ht


def cheb_dense(Ht, R, M):
    """r^T T_m(Ht) r for the rows of R, m = 0..M (plain three-term recursion)."""
    t0, t1 = R.T.copy(), Ht @ R.T
    out = [np.sum(R.T * t0, 0), np.sum(R.T * t1, 0)]
    for _ in range(2, M + 1):
        t0, t1 = t1, 2 * Ht @ t1 - t0
        out.append(np.sum(R.T * t1, 0))
    return np.array(out).T


def test_unit_vectors_give_the_exact_trace(ctx, c0):
    s, o, a, b, Ht = c0
    M = 60
    mom, tr = ctx.block_moments(M, np.eye(s.norb))
    _, tot = ctx.kpe_moments(M, L.FP64)
    assert np.abs(tr - tot).max() < 1e-10 * s.norb
    assert np.abs(mom[:, 2] - np.diag(2 * Ht @ Ht - np.eye(s.norb))).max() < 1e-12


@pytest.mark.parametrize("prec,tol", [(L.FP64, 1e-10), (L.FP32_ACC64, 2e-5)])
def test_gaussian_block_matches_dense(ctx, c0, prec, tol):
    s, o, a, b, Ht = c0
    M, nvec = 81, 24
    R = L.random_vectors(7, nvec, s.norb)
    mom, tr = ctx.block_moments(M, R, prec)
    ref = cheb_dense(Ht, R, M)
    assert np.abs(mom - ref).max() < tol * s.norb
    assert np.allclose(tr, ref.sum(0), rtol=0, atol=tol * s.norb * nvec)
    # the seeded entry point draws the same stream on the device (device and host
    # libm may differ in the last ulp), so its trace equals the host vectors' one
    st = ctx.stochastic_moments(M, nvec, seed=7, precision=prec)
    assert np.allclose(st, tr, rtol=0, atol=(1e-12 if prec == L.FP64 else tol) * s.norb * nvec)


def test_stochastic_trace_estimate(ctx, c0):
    s, o, a, b, Ht = c0
    M, nvec = 40, 512
    tr = ctx.stochastic_moments(M, nvec, seed=11) / nvec
    T = [np.eye(s.norb), Ht.copy()]
    while len(T) <= M:
        T.append(2 * Ht @ T[-1] - T[-2])
    for m, Tm in enumerate(T):
        sd = np.sqrt(2 * np.sum(Tm * Tm) / nvec)
        assert abs(tr[m] - np.trace(Tm)) < 5 * sd + 1e-9, m


def test_stochastic_diagonal(ctx, c0):
    """computeDiagonalElements in random mode: sum_r r.*(P r) ./ sum_r r.*r."""
    s, o, a, b, Ht = c0
    M, nvec = 200, 32
    ctx.kpe_moments(M, L.FP64)
    d = ctx.density(M)
    w = ctx.kpe_weights(M)
    got = ctx.density_stochastic(M, w, nvec, seed=3)
    e, V = np.linalg.eigh(Ht)
    P = (V * np.polynomial.chebyshev.chebval(e, w)) @ V.T
    R = L.random_vectors(3, nvec, s.norb)
    ref = np.sum(R * (R @ P), 0) / np.sum(R * R, 0)
    assert np.abs(got - ref).max() < 1e-9
    # the estimator is unbiased for diag(P) only as nvec grows; the sum is not exact
    assert abs(got.sum() - d["nelec"] / 2) < 0.1 * s.norb


def test_stochastic_trace_is_the_mean_of_r_P_r(ctx, c0):
    """computeAlphas' random-mode trace (withSP2_deltaH.cpp:1240-1275):
    sum_r r^T P r / nRandoms over the seeded vectors, against the dense
    formula; and with many vectors it estimates Tr P within 5 sigma
    (sigma^2 = 2 ||P||_F^2 / nvec for Gaussian vectors)."""
    s, o, a, b, Ht = c0
    M = 200
    ctx.kpe_moments(M, L.FP64)
    ctx.density(M)
    w = ctx.kpe_weights(M)
    e, V = np.linalg.eigh(Ht)
    P = (V * np.polynomial.chebyshev.chebval(e, w)) @ V.T
    for nvec, seed in ((32, 3), (500, 11)):
        got = ctx.trace_stochastic(M, w, nvec, seed=seed)
        R = L.random_vectors(seed, nvec, s.norb)
        ref = np.einsum("ij,jk,ik->", R, P, R) / nvec
        assert abs(got - ref) < 1e-9 * abs(ref)
    sigma = np.sqrt(2 * (P * P).sum() / 500)
    assert abs(got - np.trace(P)) < 5 * sigma


@pytest.mark.parametrize("prec,tol", [(L.FP64, 1e-11), (L.FP32_ACC64, 1e-5)])
def test_device_draws_match_host_stream_at_size(ctx, prec, tol):
    """18,432 orbitals: the seeded entry points (vectors drawn on the device,
    indices up to nvec * norb) against the same stream drawn on the host
    (ltb_random_vectors) through the host-vector entry points."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    M, nvec, seed = 48, 40, 2024
    R = L.random_vectors(seed, nvec, s.norb)
    _, tr = ctx.block_moments(M, R, prec)
    st = ctx.stochastic_moments(M, nvec, seed=seed, precision=prec)
    assert np.abs(st - tr).max() < tol * s.norb * nvec
    # the random-mode diagonal: device draws + device estimator vs host vectors + getKPE
    w = ctx.kpe_weights_at(M, 0.5 * (lo + hi))
    got = ctx.density_stochastic(M, w, nvec, seed=seed, precision=prec)
    Y = ctx.kpe_apply(M, w, R, prec)
    ref = np.sum(R * Y, 0) / np.sum(R * R, 0)
    assert np.abs(got - ref).max() < 10 * tol
