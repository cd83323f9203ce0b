"""HIP path against the oracle, through the C ABI (libltb.so).

Tolerances: FP64 paths agree with the FP64 oracle to rounding (stated per
test); the FP32 recursion (FP32 vectors and H~, FP64 moment accumulation) is
held to 5e-6 per atom moment and 2e-6 e per charge."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu


def pair(cells, sigma=0.05, seed=1):
    return L.System.quartz(*cells, sigma_ang=sigma, seed=seed), O.System.quartz(*cells, sigma_ang=sigma, seed=seed)


@pytest.fixture(scope="module")
def c0():
    return pair((2, 2, 2))


def oracle_atom_moments(o, a, b, nm, shift=None):
    m = o.diag_moments(a, b, nm, shift=shift)  # [orb][nm]
    return m.reshape(o.nat, 4, nm).sum(1).T     # [nm][nat]


def prepare(ctx, s, p=None):
    ctx.load(s)
    ctx.set_populations(p)
    lo, hi = ctx.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    ctx.rescale(a, b)
    return a, b


def test_gershgorin_and_shift(ctx, c0):
    s, o = c0
    ctx.load(s)
    ctx.set_populations(None)
    assert np.abs(ctx.shift()).max() == 0.0
    assert np.allclose(ctx.gershgorin(), o.gershgorin(), rtol=0, atol=1e-14)
    rng = np.random.default_rng(1)
    p = o.p0 + 0.1 * rng.standard_normal(o.nat)
    ctx.set_populations(p)
    v = o.dense_gamma() @ (p - o.p0)
    assert np.abs(ctx.shift() - v).max() < 1e-14
    assert np.allclose(ctx.gershgorin(), o.gershgorin(np.repeat(v, 4)), rtol=0, atol=1e-14)


@pytest.mark.parametrize("M", [200, 51])
def test_moments_fp64_match_oracle(ctx, c0, M):
    s, o = c0
    a, b = prepare(ctx, s)
    am, tot = ctx.kpe_moments(M, L.FP64)
    ref = oracle_atom_moments(o, a, b, M + 1)
    assert np.abs(am - ref).max() < 1e-11
    assert np.abs(tot - ref.sum(1)).max() < 1e-9
    assert np.all(am[0] == 4.0)


def test_moments_fp32_within_stated_bound(ctx, c0):
    s, o = c0
    a, b = prepare(ctx, s)
    am, _ = ctx.kpe_moments(200, L.FP32_ACC64)
    ref = oracle_atom_moments(o, a, b, 201)
    assert np.abs(am - ref).max() < 5e-6


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_ACC64])
def test_ragged_chunks_equal_single_chunk(ctx, prec):
    s = L.System.quartz(3, 3, 3, sigma_ang=0.1, seed=2)  # 972 orbitals
    prepare(ctx, s)
    try:
        ctx.set_tiling()
        ref, _ = ctx.kpe_moments(64, prec)
        ctx.set_tiling(chunk_cols=128, cols_per_lane=1, schedule=L.SCHED_BLOCKED, atoms_per_wave=1)
        got, _ = ctx.kpe_moments(64, prec)  # 8 chunks, the last one 76 columns wide
        ctx.set_tiling(chunk_cols=320, cols_per_lane=1, schedule=L.SCHED_BLOCKED, atoms_per_wave=3)
        got2, _ = ctx.kpe_moments(64, prec)
    finally:
        ctx.set_tiling()
    tol = 1e-12 if prec == L.FP64 else 2e-6
    assert np.abs(got - ref).max() < tol
    assert np.abs(got2 - ref).max() < tol


TILINGS = [dict(schedule=sch, cols_per_lane=v, nontemporal=nt, chunk_cols=ch, blocks_per_cu=bpc, order=od)
           for sch, v, nt, ch, bpc, od in [
               (L.SCHED_BLOCKED, 1, 0, 4096, 0, 0), (L.SCHED_BLOCKED, 2, 1, 2048, 0, 0),
               (L.SCHED_BLOCKED, 4, 1, 4096, 0, 0), (L.SCHED_SWEEP, 1, 1, 4096, 4, 1),
               (L.SCHED_SWEEP, 2, 0, 4096, 2, 2), (L.SCHED_SWEEP, 4, 1, 8192, 8, 1),
               (L.SCHED_SWEEP, 4, 1, 4096, 2, 2), (L.SCHED_SWEEP, 1, 1, 5000, 3, 1)]]


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_ACC64])
def test_every_tiling_gives_the_same_moments(ctx, prec):
    """18,432 orbitals: both schedules, V in {1,2,4}, cached / non-temporal,
    sweep in spatial-strip and in atom-index order, chunk widths that leave a
    ragged last chunk (5000 -> 3 x 5056 + 3264)."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    prepare(ctx, s)
    try:
        ctx.set_tiling(schedule=L.SCHED_BLOCKED, cols_per_lane=1, chunk_cols=4096)
        ref, _ = ctx.kpe_moments(24, L.FP64)
        for t in TILINGS:
            ctx.set_tiling(**t)
            got, _ = ctx.kpe_moments(24, prec)
            tol = 1e-12 if prec == L.FP64 else 3e-6
            assert np.abs(got - ref).max() < tol, t
    finally:
        ctx.set_tiling()


def quartz_cluster(cells, sigma=0.05, seed=1):
    """the periodic cell's atoms as a finite cluster (no images)"""
    z, xyz = L.System.quartz(*cells, sigma_ang=sigma, seed=seed).coords()
    return L.System.from_coords(z, xyz, None, periodic_mask=0)


CONES = [  # system, tiling (sweep schedule, atom-index order)
    (lambda: L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1), dict(chunk_cols=4096, cols_per_lane=4)),
    (lambda: L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1), dict(chunk_cols=1024, cols_per_lane=1)),
    # 1728 orbitals in chunks of 1024: the first chunk holds more than half the atoms
    (lambda: L.System.quartz(4, 4, 3, sigma_ang=0.1, seed=3), dict(chunk_cols=1024, cols_per_lane=1)),
    (lambda: quartz_cluster((3, 3, 3)), dict(chunk_cols=512, cols_per_lane=1)),
]


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_ACC64])
@pytest.mark.parametrize("case", range(len(CONES)))
def test_light_cone_equals_the_full_sweep(ctx, case, prec):
    """Sweep steps restricted to each chunk's light cone give the full sweep's
    moments (the vectors are the same; only the order of the FP64 dot sums
    changes) for low orders, where the cones are narrow, and for M = 512,
    where they fill the cell; the counted bytes shrink accordingly."""
    make, tiling = CONES[case]
    s = make()
    prepare(ctx, s)
    try:
        ctx.set_tiling(schedule=L.SCHED_SWEEP, blocks_per_cu=2, nontemporal=1, order=2, **tiling)
        for M in (32, 512):
            ctx.set_timing(True)
            ctx.set_light_cone(False)
            full, tfull = ctx.kpe_moments(M, prec)
            bfull = ctx.kpe_stats()["bytes"]
            ctx.set_light_cone(True)
            cone, tcone = ctx.kpe_moments(M, prec)
            bcone = ctx.kpe_stats()["bytes"]
            assert np.abs(cone - full).max() < 1e-12 * max(1.0, np.abs(full).max()), M
            assert np.abs(tcone - tfull).max() < 1e-12 * np.abs(tfull).max(), M
            assert bcone <= bfull
            if M == 32 and s.nat > 1000:
                assert bcone < 0.95 * bfull
    finally:
        ctx.set_light_cone(True)
        ctx.set_timing(False)
        ctx.set_tiling()


def test_autotune_picks_a_valid_tiling(ctx):
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    prepare(ctx, s)
    try:
        best64, _ = ctx.autotune(L.FP64, trial_steps=2)
        best, gbs = ctx.autotune(L.FP32_ACC64, trial_steps=2)
        assert gbs > 100
        assert best["cols_per_lane"] in (1, 2, 4) and best["schedule"] in (1, 2)
        # each precision keeps its own tiling
        assert ctx.get_tiling(L.FP64) == best64 and ctx.get_tiling(L.FP32_ACC64) == best
        am, _ = ctx.kpe_moments(24, L.FP32_ACC64)
        ctx.set_tiling(schedule=L.SCHED_BLOCKED, cols_per_lane=1)
        ref, _ = ctx.kpe_moments(24, L.FP64)
        assert np.abs(am - ref).max() < 3e-6
    finally:
        ctx.set_tiling()


def test_density_matches_oracle(ctx, c0):
    s, o = c0
    prepare(ctx, s)
    ctx.kpe_moments(200, L.FP64)
    d = ctx.density(200)
    r = o.density(200)
    assert d["nelec"] == pytest.approx(384.0, abs=1e-8)
    assert np.abs(d["pop"] - r["pop"]).max() < 1e-9
    assert d["eband"] == pytest.approx(r["eband"], abs=1e-8)
    assert abs(d["mu"] - r["mu"]) < 1e-6  # mu in a 6.6 eV gap is weakly determined


def test_density_step_function_mode(ctx, c0):
    s, o = c0
    prepare(ctx, s)
    ctx.kpe_moments(200, L.FP64)
    d = ctx.density(200, kT_ev=0.0)
    r = o.density(200, kT_ev=0.0)
    assert np.abs(d["pop"] - r["pop"]).max() < 1e-9


def test_kpe_apply_matches_getKPE(ctx, c0):
    s, o = c0
    a, b = prepare(ctx, s)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((3, s.norb))
    c = O.fd_coeffs(200, -0.2, 1e-3, 8192)
    w = O.kpe_weights(c, O.gmM(200))
    Y = ctx.kpe_apply(200, w, X, L.FP64)
    for j in range(3):
        ref = o.kpe_apply(a, b, w, X[j])
        assert np.abs(Y[j] - ref).max() < 1e-11
    Y32 = ctx.kpe_apply(200, w, X, L.FP32_ACC64)
    assert np.abs(Y32 - Y).max() < 1e-4


def test_scc_fp64_matches_oracle(ctx, c0):
    s, o = c0
    ctx.load(s)
    g = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    r = o.scc(200, tol=1e-6, maxit=300)
    assert g["status"] == 0 and r["status"] == 0
    assert abs(g["iters"] - r["iters"]) <= 1
    assert np.abs(g["q"] - r["q"]).max() < 1e-8
    assert g["eband"] == pytest.approx(r["eband"], abs=1e-7)
    assert 2 * g["pop"].sum() == pytest.approx(384.0, abs=1e-8)


EV = L.HARTREE_EV


@pytest.fixture(scope="module")
def c0_fixed_point(c0):
    """Oracle SCC converged to 1e-12 e and E_band evaluated at q*."""
    _, o = c0
    r = o.scc(200, tol=1e-12, maxit=400)
    v = o.dense_gamma() @ (r["pop"] - o.p0)
    d = o.density(200, shift=np.repeat(v, 4))
    return r["q"], d["eband"]


def test_scc_mixed_precision_matches_fp64(ctx, c0, c0_fixed_point):
    """FP32 bulk, FP64 iterations once |dq| < 1e-6 until an FP64 iteration
    converges: charges within 1e-6 e and E_band (fixed-point estimate) within
    1e-6 eV/atom of the plain FP64 solve."""
    s, o = c0
    qstar, estar = c0_fixed_point
    ctx.load(s)
    ref = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    mix = ctx.scc(200, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=300)
    assert mix["status"] == 0 and 1 <= mix["fp64_iters"] <= 5
    assert np.abs(mix["q"] - ref["q"]).max() < 1e-6
    assert abs(mix["eband_fp"] - ref["eband_fp"]) * EV / s.nat < 1e-6
    # both against the tightly converged oracle fixed point
    assert np.abs(ref["q"] - qstar).max() < 1e-7
    assert abs(ref["eband_fp"] - estar) * EV / s.nat < 2e-6
    # the uncorrected last-iteration value is first order in |q_in - q*|
    assert abs(ref["eband"] - estar) * EV / s.nat > 1e-5


def test_scc_with_hubbard_onsite_gamma_matches_oracle(ctx):
    """The +U variant (Gamma_AA = U_Si, U_O from the reference's unused
    Hubbard parameters) through ltb_set_gamma_onsite: the device Gamma holds U
    on its diagonal, the FP64 SCC equals the oracle's with the same Gamma
    (1e-8 e), and reloading the system restores the reference's -1."""
    s, o = pair((2, 2, 2))
    ctx.load(s)
    z, _ = s.coords()
    u = L.hubbard_u(z)
    ctx.set_gamma_onsite(u)
    rp, col, val = ctx.gamma()
    diag = np.array([val[k] for i in range(s.nat) for k in range(rp[i], rp[i + 1]) if col[k] == i])
    assert np.array_equal(diag, u)
    r = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=600)
    o.set_gamma_onsite(u)
    ref = o.scc(200, tol=1e-6, maxit=600)
    assert ref["status"] == 0 and abs(r["iters"] - ref["iters"]) <= 1
    assert np.abs(r["q"] - ref["q"]).max() < 1e-8
    ctx.load(s)
    rp, col, val = ctx.gamma()
    assert all(val[k] == -1.0 for i in range(s.nat) for k in range(rp[i], rp[i + 1]) if col[k] == i)
    with pytest.raises(ValueError):
        ctx.set_gamma_onsite(u[:-1])


def test_scc_graph_replays_match_eager(ctx, c0):
    """hipGraph-captured iterations (replayed many times, host reads between
    replays) reproduce the eager loop bit for bit."""
    s, _ = c0
    ctx.load(s)
    for prec, sw in [(L.FP64, 0.0), (L.FP32_ACC64, 1e-4)]:
        eager = ctx.scc(200, precision=prec, switch_tol=sw, tol=1e-6, max_iter=300, use_graph=0)
        graph = ctx.scc(200, precision=prec, switch_tol=sw, tol=1e-6, max_iter=300, use_graph=1)
        assert graph["graph_launches"] >= graph["iters"] - 3
        assert graph["iters"] == eager["iters"]
        assert np.array_equal(graph["q"], eager["q"])
        assert np.array_equal(graph["hist"], eager["hist"])


def test_scc_resumes_exactly_from_mixed_populations(ctx, c0):
    """An SCC cut after 10 iterations and restarted from the mixed populations
    repeats the uninterrupted run bit for bit (linear mixing has no other state)."""
    s, _ = c0
    ctx.load(s)
    full = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    first = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=10, check=False)
    assert first["status"] == L.LTB_NOT_CONVERGED
    rest = ctx.scc(200, guess=ctx.mixed_populations(), precision=L.FP64, tol=1e-6, max_iter=300)
    assert rest["iters"] == full["iters"] - 10
    assert np.array_equal(rest["q"], full["q"])
    assert np.array_equal(np.concatenate([first["hist"], rest["hist"]]), full["hist"])


def test_scc_iteration_cap_reports_not_converged(ctx, c0):
    s, _ = c0
    ctx.load(s)
    r = ctx.scc(200, precision=L.FP64, max_iter=3, check=False)
    assert r["status"] == L.LTB_NOT_CONVERGED and r["iters"] == 3


def test_nonfinite_populations_are_reported(ctx, c0):
    s, o = c0
    ctx.load(s)
    p = o.p0.copy()
    p[5] = np.nan
    ctx.set_populations(p)
    with pytest.raises(L.LtbError) as e:
        ctx.gershgorin()
    assert e.value.status == L.LTB_ERR_NONFINITE
    ctx.set_populations(o.p0)


def test_largest_cell_properties(ctx):
    """27,648 atoms (110,592 orbitals, the largest single-GPU cell): many chunks in
    both precisions, sum rules, and FP32 against FP64 per atom moment."""
    s = L.System.quartz(16, 16, 12, sigma_ang=0.08, seed=1)
    prepare(ctx, s)
    M = 16
    am64, tot64 = ctx.kpe_moments(M, L.FP64)
    am32, tot32 = ctx.kpe_moments(M, L.FP32_ACC64)
    assert tot64[0] == s.norb and np.all(am64[0] == 4.0)
    assert np.all(np.abs(am64) <= 4 + 1e-9)           # |[T_m]_jj| <= 1 for a spectrum inside [-1, 1]
    assert np.abs(am32 - am64).max() < 5e-6
    assert np.abs(tot64 - am64.sum(1)).max() < 1e-8 * s.norb


def test_c1_size_properties(ctx):
    """4,608 atoms (18,432 orbitals): FP32 vs FP64 and sum rules at full size."""
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=1)
    prepare(ctx, s)
    M = 64
    am64, tot64 = ctx.kpe_moments(M, L.FP64)
    am32, tot32 = ctx.kpe_moments(M, L.FP32_ACC64)
    assert tot64[0] == s.norb
    assert np.all(np.abs(am64) <= 4 + 1e-9)          # |[T_m]_jj| <= 1
    assert np.abs(am32 - am64).max() < 5e-6
    d = ctx.density(M)
    assert 2 * d["pop"].sum() == pytest.approx(s.nelectrons, rel=1e-12)
    assert d["nelec"] == pytest.approx(s.nelectrons, rel=1e-12)
