"""The two-level SCC of ltb_scc_run (M_lo > 0) through the C ABI.

Most iterations evaluate the density at order M_lo plus the defect
F_M(q_ref) - F_M_lo(q_ref) of the last full-order iteration; convergence is
declared on full-order iterations only, so the answer is the plain SCC's
fixed point (the loop of withSP2_deltaH.cpp:1815-1840 with linear charge
mixing in place of updateHtotal).

* 72 atoms (2x2x2, sigma 0.05 A, seed 1), M = 200, FP64 throughout: every
  iteration's |dq| equals the oracle's two-level loop (same phases, same
  switches) to 1e-9 e, and the charges equal the oracle's plain SCC fixed
  point to 1e-7 e; hipGraph replays of the low-order iteration equal eager
  iterations bit for bit.
* 4,608 atoms (8x8x8, sigma 0.05 A, seed 1), M = 512: FP32 low-order and
  early full-order iterations, FP64 full-order iterations once |dq| < 1e-2 e,
  from the neutral start, against the plain FP64 solve: every charge within
  1e-6 e and the fixed-point band energy within 1e-6 eV/atom.
"""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu
EV = L.HARTREE_EV


@pytest.fixture(scope="module")
def c0():
    return L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1), O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)


def test_two_level_trajectory_matches_the_oracle(ctx, c0):
    s, o = c0
    ctx.load(s)
    r = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=1000, use_graph=1, M_lo=32)
    ref = o.scc_two_level(200, 32, tol=1e-6)
    fixed = o.scc(200, tol=1e-10, maxit=600)
    print(f"72 atoms two-level: {r['iters']} iterations ({r['lo_iters']} at M=32, {r['graph_launches']} graph "
          f"replays), oracle {ref['iters']} ({ref['lo_iters']}); max |dq - dq_oracle| "
          f"{np.abs(r['hist'] - ref['hist']).max() if r['iters'] == ref['iters'] else float('nan'):.2e}")
    assert r["status"] == 0 and ref["status"] == 0
    assert r["iters"] == ref["iters"] and r["lo_iters"] == ref["lo_iters"]
    assert r["fp64_iters"] == r["iters"] - r["lo_iters"]
    assert np.abs(r["hist"] - ref["hist"]).max() < 1e-9
    assert np.abs(r["q"] - ref["q"]).max() < 1e-9
    assert np.abs(r["q"] - fixed["q"]).max() < 1e-7
    assert r["graph_launches"] > 0 and r["graph_launches"] < r["lo_iters"]
    assert abs(r["eband"] - ref["eband"]) * EV / s.nat < 1e-9


def test_two_level_graph_equals_eager(ctx, c0):
    s, _ = c0
    ctx.load(s)
    kw = dict(precision=L.FP64, tol=1e-6, max_iter=1000, M_lo=32)
    g = ctx.scc(200, use_graph=1, **kw)
    e = ctx.scc(200, use_graph=0, **kw)
    assert g["iters"] == e["iters"] and g["graph_launches"] > 0 and e["graph_launches"] == 0
    assert np.array_equal(g["q"], e["q"]) and np.array_equal(g["hist"], e["hist"])
    assert g["eband_fp"] == e["eband_fp"] and g["mu_ha"] == e["mu_ha"]


def test_two_level_mixed_against_plain_fp64_at_4608_atoms(ctx):
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    for prec in (L.FP64, L.FP32_ACC64):
        ctx.autotune(prec, 3)
    ref = ctx.scc(512, precision=L.FP64, tol=1e-6, max_iter=200, use_graph=1)
    r = ctx.scc(512, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000, use_graph=1, M_lo=32)
    dq = np.abs(r["q"] - ref["q"]).max()
    deb = abs(r["eband_fp"] - ref["eband_fp"]) * EV / s.nat
    full = r["iters"] - r["lo_iters"]
    print(f"4,608 atoms: plain FP64 {ref['iters']} iterations {ref['seconds']:.2f} s; two-level {full} full-order "
          f"({r['fp64_iters']} FP64) + {r['lo_iters']} at M=32, {r['seconds']:.2f} s: max |dq| {dq:.2e} e, "
          f"|dE_band| {deb:.2e} eV/atom, |dmu| {abs(r['mu_ha'] - ref['mu_ha']) * EV:.2e} eV")
    assert ref["status"] == 0 and r["status"] == 0
    assert dq < 1e-6 and deb < 1e-6
    assert 1 <= r["fp64_iters"] <= full <= 8
    assert r["q"].sum() == pytest.approx(s.nelectrons, abs=1e-6)


def test_two_level_parameter_checks(ctx, c0):
    s, _ = c0
    ctx.load(s)
    for kw in (dict(M_lo=200), dict(M_lo=1), dict(M_lo=-1), dict(M_lo=32, lo_tol=-1.0)):
        with pytest.raises(L.LtbError) as e:
            ctx.scc(200, precision=L.FP64, **kw)
        assert e.value.status == L.LTB_ERR_ARG
    b = L.Batch(ctx, [s])
    try:
        prm = L.SccParams(200, 0.05 / EV, 8192, 0.2, 1e-6, 10, L.FP64, 0.0, 0, 0, 32, 0.0)
        z = np.zeros(s.nat)
        zi = np.zeros(1, np.int32)
        zd = np.zeros(1)
        st = L.lib().ltb_batch_scc_run(b._h, L.C.byref(prm), None, L._d(z), L._i(zi), L._d(zd), L._d(zd),
                                       L._i(zi), None)
        assert st == L.LTB_ERR_ARG
    finally:
        b.close()
