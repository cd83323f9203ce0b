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
* Batches (ltb_batch_scc_run with M_lo > 0): five uneven cells and the
  256-cell workload against each cell's own plain FP64 solve.
* 15,552 atoms (12x12x12, sigma 0.10 A, M = 512) and 27,648 atoms (16x16x12,
  sigma 0.08 A, M = 1024): the shipped learned guess plus the two-level SCC,
  then one plain FP64 iteration from the answer (the plain solves take 5 and
  33 minutes, too long to repeat here).
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


def test_two_level_cap_and_converged_start(ctx, c0):
    """The iteration cap counts both kinds of iteration and reports
    NOT_CONVERGED; started from its own answer, the two-level SCC stops after
    its first (full-order) iteration with the same charges; FP32 throughout
    (switch_tol 0) converges on an FP32 full-order iteration within the FP32
    rounding of the FP64 answer."""
    s, _ = c0
    ctx.load(s)
    r = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=7, M_lo=32, check=False)
    assert r["status"] == L.LTB_NOT_CONVERGED and r["iters"] == 7 and r["lo_iters"] == 6
    ref = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=1000, M_lo=32)
    again = ctx.scc(200, guess=ref["pop"], precision=L.FP64, tol=1e-6, max_iter=1000, M_lo=32)
    assert again["status"] == 0 and again["iters"] == 1 and again["lo_iters"] == 0
    assert np.abs(again["q"] - ref["q"]).max() < 1e-7
    f32 = ctx.scc(200, precision=L.FP32_ACC64, switch_tol=0.0, tol=1e-6, max_iter=1000, M_lo=32, use_graph=1)
    assert f32["status"] == 0 and f32["fp64_iters"] == 0
    assert np.abs(f32["q"] - ref["q"]).max() < 2e-6


def test_two_level_parameter_checks(ctx, c0):
    s, _ = c0
    ctx.load(s)
    for kw in (dict(M_lo=200), dict(M_lo=1), dict(M_lo=-1), dict(M_lo=32, lo_tol=-1.0)):
        with pytest.raises(L.LtbError) as e:
            ctx.scc(200, precision=L.FP64, **kw)
        assert e.value.status == L.LTB_ERR_ARG
    b = L.Batch(ctx, [s])
    try:
        for kw in (dict(M_lo=200), dict(M_lo=1), dict(M_lo=32, lo_tol=-1.0)):
            with pytest.raises(L.LtbError) as e:
                b.scc(200, precision=L.FP32_ACC64, switch_tol=1e-6, **kw)
            assert e.value.status == L.LTB_ERR_ARG
    finally:
        b.close()


def test_two_level_low_order_chunk_width(ctx):
    """tiling.lo_chunk_cols: the low-order recursions of the two-level SCC at
    half the chunk width (narrower light cones) follow the run at the full
    width up to the order of the FP64 dot sums.  4,608 atoms (18,432
    orbitals), FP64 throughout, 2048-column chunks, low-order chunks of 2048 or
    1024 columns: the first 12 iterations (1 full-order, 11 at M_lo = 32)
    agree to 1e-12 e in |dq| and in the charges.  The autotuner picks either
    width and reports it."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    for prec in (L.FP64, L.FP32_ACC64):
        t, _ = ctx.autotune(prec, 3)
        assert t["lo_chunk_cols"] in (t["chunk_cols"], t["chunk_cols"] // 2), t
        assert ctx.get_tiling(prec) == t
    runs = {}
    try:
        for lo_cols in (0, 1024):
            ctx.set_tiling(chunk_cols=2048, cols_per_lane=2, schedule=L.SCHED_SWEEP, blocks_per_cu=2, nontemporal=1,
                           order=2, lo_chunk_cols=lo_cols)
            assert ctx.get_tiling(L.FP64)["lo_chunk_cols"] == (lo_cols or 2048)
            runs[lo_cols] = ctx.scc(512, precision=L.FP64, tol=1e-6, max_iter=12, use_graph=1, M_lo=32, check=False)
    finally:
        ctx.set_tiling()
    a, b = runs[0], runs[1024]
    print(f"4,608 atoms FP64, low-order chunks 2048 vs 1024 columns: max |d hist| "
          f"{np.abs(a['hist'] - b['hist']).max():.2e}, max |dq| {np.abs(a['q'] - b['q']).max():.2e} e")
    for r in (a, b):
        assert r["status"] == L.LTB_NOT_CONVERGED and r["iters"] == 12 and r["lo_iters"] == 11
    assert np.abs(a["hist"] - b["hist"]).max() < 1e-12
    assert np.abs(a["q"] - b["q"]).max() < 1e-12


# uneven cells (L, M, N, sigma A, seed), 72 to 162 atoms
CELLS = [(2, 2, 2, 0.05, 1), (3, 2, 2, 0.10, 7), (2, 3, 2, 0.02, 11), (2, 2, 3, 0.15, 3),
         (3, 3, 2, 0.08, 5)]


def test_batch_two_level_matches_single_fp64_solves(ctx):
    """Batched two-level SCC (FP32 structures: low-order iterations at M_lo =
    32 between full-order ones, then plain FP64 iterations once a full-order
    |dq| < 1e-6 e): every structure within 1e-6 e and 1e-6 eV/atom of its own
    plain FP64 solve; graph replays equal eager iterations bit for bit."""
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in CELLS]
    singles = []
    for s in systems:
        ctx.load(s)
        singles.append(ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300))
    with L.Batch(ctx, systems) as b:
        kw = dict(precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=2000, M_lo=32)
        r = b.scc(200, use_graph=1, **kw)
        e = b.scc(200, use_graph=0, **kw)
    print(f"batch two-level: {r['batch_iters']} batch iterations ({r['lo_iters']} low-order only, "
          f"{r['fp64_iters']} with FP64), per structure {list(r['iters'])}")
    assert r["status"] == 0 and (r["struct_status"] == 0).all()
    assert r["lo_iters"] > 0 and r["fp64_iters"] >= 1 and r["graph_launches"] > 0
    for k, (q, g) in enumerate(zip(b.split(r["q"]), singles)):
        assert np.abs(q - g["q"]).max() < 1e-6, k
        assert abs(r["eband_fp"][k] - g["eband_fp"]) * EV / systems[k].nat < 1e-6, k
        assert q.sum() == pytest.approx(systems[k].nelectrons, abs=1e-6)
    assert np.array_equal(r["iters"], e["iters"]) and np.array_equal(r["q"], e["q"])
    assert np.array_equal(r["eband_fp"], e["eband_fp"])


def test_batch_two_level_of_256_cells(ctx):
    """The 256-cell batched workload (86,211 atoms, M = 256) with the two-level
    SCC: every cell converges and conserves its electrons; eight cells across
    the size range match their own plain FP64 solves to 1e-6 e and 1e-6
    eV/atom."""
    from largetb_amd.learned import batch_structures

    cells = batch_structures(256, seed=1)
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in cells]
    with L.Batch(ctx, systems) as b:
        r = b.scc(256, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=2000, use_graph=1, M_lo=32)
        qs = b.split(r["q"])
    print(f"256 cells two-level: {r['batch_iters']} batch iterations ({r['lo_iters']} low-order only), "
          f"{r['seconds']:.2f} s")
    assert r["status"] == 0 and (r["struct_status"] == 0).all()
    for k, s in enumerate(systems):
        assert qs[k].sum() == pytest.approx(s.nelectrons, abs=1e-6)
    sizes = np.array([s.nat for s in systems])
    order = np.argsort(sizes, kind="stable")
    for k in order[np.linspace(0, len(order) - 1, 8).round().astype(int)]:
        ctx.load(systems[k])
        g = ctx.scc(256, precision=L.FP64, tol=1e-6, max_iter=400, use_graph=1)
        assert np.abs(qs[k] - g["q"]).max() < 1e-6, k
        assert abs(r["eband_fp"][k] - g["eband_fp"]) * EV / systems[k].nat < 1e-6, k


def learned_two_level_then_plain_iteration(ctx, s, M):
    """The shipped learned guess, the two-level SCC as the benchmark runs it
    (M_lo = 32; FP32 full-order iterations until |dq| < 1e-2 e, FP64 after),
    then one plain FP64 iteration from its answer."""
    from largetb_amd import learned
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    for prec in (L.FP64, L.FP32_ACC64):
        ctx.autotune(prec, 3)
    m, _ = learned.load(ctx)
    guess = m.predict(s)
    m.close()
    r = ctx.scc(M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000, use_graph=1,
                M_lo=32)
    chk = ctx.scc(M, guess=r["pop"], precision=L.FP64, tol=1e-30, max_iter=1, check=False)
    full = r["iters"] - r["lo_iters"]
    deb = abs(chk["eband_fp"] - r["eband_fp"]) * EV / s.nat
    print(f"{s.nat} atoms, M={M}: learned guess off by {np.abs(2 * guess - r['q']).max():.2e} e; two-level {full} "
          f"full-order ({r['fp64_iters']} FP64) + {r['lo_iters']} low-order, {r['seconds']:.2f} s; plain FP64 "
          f"iteration from the answer: |dq| {chk['hist'][0]:.2e} e, |dE_band| {deb:.2e} eV/atom")
    return r, chk, full, deb


def test_c3_learned_two_level_reaches_the_plain_fixed_point(ctx):
    """The headline workload (12x12x12, 15,552 atoms, sigma 0.10 A, seed 1,
    M = 512) as the benchmark runs it converges in at most 4 full-order
    iterations.  One plain FP64 iteration started from its answer moves every
    charge by < 1e-7 e and the band energy by < 1e-7 eV/atom, so the answer is
    the fixed point the plain FP64 solve converges to (that solve stops within
    ~1e-8 e of it)."""
    s = L.System.quartz(12, 12, 12, sigma_ang=0.10, seed=1)
    r, chk, full, deb = learned_two_level_then_plain_iteration(ctx, s, 512)
    assert r["status"] == 0 and full <= 4 and r["fp64_iters"] >= 1
    assert chk["hist"][0] < 1e-7 and deb < 1e-7
    assert r["q"].sum() == pytest.approx(s.nelectrons, abs=1e-6)


def test_c4_learned_two_level_reaches_the_plain_fixed_point(ctx):
    """The largest single-GPU cell (16x16x12, 27,648 atoms, 110,592 orbitals,
    sigma 0.08 A, seed 1) at its configured M = 1024, the full pipeline:
    learned guess and two-level SCC in at most 4 full-order iterations, then
    one plain FP64 iteration from the answer moves every charge by < 1e-7 e and
    the band energy by < 1e-7 eV/atom."""
    s = L.System.quartz(16, 16, 12, sigma_ang=0.08, seed=1)
    r, chk, full, deb = learned_two_level_then_plain_iteration(ctx, s, 1024)
    assert r["status"] == 0 and full <= 4 and r["fp64_iters"] >= 1
    assert chk["hist"][0] < 1e-7 and deb < 1e-7
    assert r["q"].sum() == pytest.approx(s.nelectrons, abs=1e-6)
