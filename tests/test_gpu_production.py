"""The production configurations through the C ABI, at full size.

* The 15,552-atom 12x12x12 cell (sigma 0.10 A, seed 1) and the 4,608-atom
  8x8x8 cell (sigma 0.05 A, seed 1) at M = 512 with the autotuned tilings of
  both precisions (the tilings the benchmark and the SCC runs use): per-orbital
  moments [T_m]_jj of sampled unit-vector columns of every chunk -- chunk
  edges, lane and column-group edges, the ragged last chunk -- against the
  oracle's getKPE op sequence (withSP2_deltaH.cpp:1148-1168 on e_j, the exact
  mode of :1361-1371).
* The narrow-last-chunk paths of the sweep schedule (fallback to the blocked
  schedule, and its partial-row clamp).
* The 12x12x12 density and two SCC iterations in both precisions.
* The 256-cell batch (2x2x2 .. 5x5x4, seed 1) at full size.
* The largest single-GPU cell (16x16x12, 27,648 atoms) at M = 1024 in both
  precisions.

Tolerances: FP64 moments agree with the FP64 oracle to 1e-11 (summation
order only); FP32 recursion (FP32 vectors and H~, FP64 accumulation) to
5e-6 per orbital moment; FP32 vs FP64 charges to 2e-6 e and band energy to
1e-6 eV/atom."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu
EV = L.HARTREE_EV
M = 512
TOL = {L.FP64: 1e-11, L.FP32_ACC64: 5e-6}


def prepare(ctx, s):
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    ctx.rescale(a, b)
    return a, b


def chunk_samples(norb, ld, V, rng, extra=2):
    """columns of every chunk [j0, j1): its first and last columns, the last
    column of the first lane and of the first column group, the first column
    of the second group, and `extra` random ones"""
    W = 64 * V
    cols = set()
    for j0 in range(0, norb, ld):
        j1 = min(j0 + ld, norb)
        for c in (j0, j0 + V - 1, j0 + W - 1, j0 + W, j1 - 1):
            if j0 <= c < j1:
                cols.add(c)
        cols.update(int(x) for x in rng.integers(j0, j1, extra))
    return np.array(sorted(cols), np.int32)


def product_moments(ctx, cols):
    """[len(cols), M+1] per-orbital moments of the last ltb_kpe_moments call"""
    out = np.zeros((len(cols), M + 1))
    runs = np.split(cols, np.nonzero(np.diff(cols) != 1)[0] + 1)
    k = 0
    for r in runs:
        out[k:k + len(r)] = ctx.orbital_moments(M, int(r[0]), int(r[-1]) + 1).T
        k += len(r)
    return out


class Production:
    """A cell with both precisions' autotuned tilings, the moments of each and
    the oracle's moments of the union of the sampled columns."""

    def __init__(self, ctx, cells, sigma):
        self.s = L.System.quartz(*cells, sigma_ang=sigma, seed=1)
        self.o = O.System.quartz(*cells, sigma_ang=sigma, seed=1)
        a, b = prepare(ctx, self.s)
        self.a, self.b = a, b
        rng = np.random.default_rng(7)
        self.tiling, self.cols, self.mom, self.atoms, self.dens = {}, {}, {}, {}, {}
        ctx.keep_orbital_moments(True)
        try:
            for prec in (L.FP64, L.FP32_ACC64):
                t, _ = ctx.autotune(prec, 3)
                self.tiling[prec] = t
                self.atoms[prec], _ = ctx.kpe_moments(M, prec)
                self.dens[prec] = ctx.density(M)
                self.cols[prec] = chunk_samples(self.s.norb, t["chunk_cols"], t["cols_per_lane"], rng)
                self.mom[prec] = product_moments(ctx, self.cols[prec])
        finally:
            ctx.keep_orbital_moments(False)
        union = np.union1d(self.cols[L.FP64], self.cols[L.FP32_ACC64]).astype(np.int32)
        ref = self.o.diag_moments_at(a, b, M + 1, union)
        self.ref = {p: ref[np.searchsorted(union, self.cols[p])] for p in self.cols}


@pytest.fixture(scope="module")
def c3(ctx):
    return Production(ctx, (12, 12, 12), 0.10)


@pytest.fixture(scope="module")
def c1(ctx):
    return Production(ctx, (8, 8, 8), 0.05)


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_ACC64])
@pytest.mark.parametrize("cell", ["c1", "c3"])
def test_production_tiling_matches_oracle_on_sampled_columns(request, cell, prec):
    p = request.getfixturevalue(cell)
    t = p.tiling[prec]
    err = np.abs(p.mom[prec] - p.ref[prec]).max()
    print(f"{cell} prec {prec} tiling {t}: {len(p.cols[prec])} columns, max |dmoment| = {err:.3e}")
    assert t["schedule"] in (L.SCHED_BLOCKED, L.SCHED_SWEEP)
    # every chunk is sampled, the ragged last one included
    nchunks = -(-p.s.norb // t["chunk_cols"])
    assert len(np.unique(p.cols[prec] // t["chunk_cols"])) == nchunks
    assert np.all(p.mom[prec][:, 0] == 1.0)
    assert err < TOL[prec]


@pytest.mark.parametrize("cell", ["c1", "c3"])
def test_production_fp32_against_fp64(request, cell):
    """Every per-atom moment, the charges, band energy and electron count of
    the neutral-start density in the FP32 form against FP64."""
    p = request.getfixturevalue(cell)
    a64, a32 = p.atoms[L.FP64], p.atoms[L.FP32_ACC64]
    d64, d32 = p.dens[L.FP64], p.dens[L.FP32_ACC64]
    dmom = np.abs(a32 - a64).max()
    dq = 2 * np.abs(d32["pop"] - d64["pop"]).max()
    deb = abs(d32["eband"] - d64["eband"]) * EV / p.s.nat
    print(f"{cell}: max |dmoment| per atom {dmom:.3e}, max |dq| {dq:.3e} e, |dE_band| {deb:.3e} eV/atom")
    assert np.all(a64[0] == 4.0) and np.all(np.abs(a64) <= 4 + 1e-9)
    assert dmom < 2e-5
    assert dq < 2e-6
    assert deb < 1e-6
    for d in (d32, d64):
        assert d["nelec"] == pytest.approx(p.s.nelectrons, rel=1e-12)


def test_c1_scc_mixed_precision_against_plain_fp64(ctx, c1):
    """The whole 4,608-atom SCC to |dq| < 1e-6 e (hipGraph iterations): FP32
    bulk with the FP64 finish against the plain FP64 solve -- every charge
    within 1e-6 e, the fixed-point band energy within 1e-6 eV/atom, and both
    conserve the electrons."""
    ctx.load(c1.s)
    ref = ctx.scc(M, precision=L.FP64, tol=1e-6, max_iter=200, use_graph=1)
    mix = ctx.scc(M, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=200, use_graph=1)
    dq = np.abs(mix["q"] - ref["q"]).max()
    deb = abs(mix["eband_fp"] - ref["eband_fp"]) * EV / c1.s.nat
    print(f"c1 SCC: plain {ref['iters']} iterations, mixed {mix['iters']} ({mix['fp64_iters']} FP64): "
          f"max |dq| {dq:.3e} e, |dE_band| {deb:.3e} eV/atom")
    assert ref["status"] == 0 and mix["status"] == 0 and 1 <= mix["fp64_iters"] <= 5
    assert dq < 1e-6 and deb < 1e-6
    for r in (ref, mix):
        assert r["q"].sum() == pytest.approx(c1.s.nelectrons, abs=1e-6)


def test_c3_scc_iterations_fp32_against_fp64(ctx, c3):
    """Two SCC iterations from the neutral start in each precision: the
    mixed charges and the charges of the second density agree to 2e-6 e."""
    ctx.load(c3.s)
    r64 = ctx.scc(M, precision=L.FP64, max_iter=2, tol=1e-30, check=False, use_graph=1)
    p64 = ctx.mixed_populations()
    r32 = ctx.scc(M, precision=L.FP32_ACC64, max_iter=2, tol=1e-30, check=False, use_graph=1)
    p32 = ctx.mixed_populations()
    assert r64["status"] == L.LTB_NOT_CONVERGED and r32["status"] == L.LTB_NOT_CONVERGED
    assert r64["iters"] == r32["iters"] == 2
    assert np.abs(r32["q"] - r64["q"]).max() < 2e-6
    assert 2 * np.abs(p32 - p64).max() < 2e-6
    assert abs(r32["eband"] - r64["eband"]) * EV / c3.s.nat < 1e-6
    assert r32["hist"][0] == pytest.approx(r64["hist"][0], abs=2e-6)


NARROW = [
    # 3x3x3 (972 orbitals), sweep, V=1, 640-column chunks: the last chunk has
    # 332 columns = 6 groups of 64, fewer than the 8 the sweep needs -> blocked
    ((3, 3, 3), 0.1, dict(chunk_cols=640, cols_per_lane=1, schedule=L.SCHED_SWEEP, blocks_per_cu=2)),
    # 8x8x8 (18,432 orbitals), sweep, V=4, 3584-column chunks, 1 workgroup per
    # CU: the last chunk has 512 columns = 2 groups -> blocked, with more
    # atom ranges than the sweep's partial rows -> clamp
    ((8, 8, 8), 0.05, dict(chunk_cols=3584, cols_per_lane=4, schedule=L.SCHED_SWEEP, blocks_per_cu=1)),
]


@pytest.mark.parametrize("prec", [L.FP64, L.FP32_ACC64])
@pytest.mark.parametrize("cells,sigma,tiling", NARROW)
def test_narrow_last_chunk_of_the_sweep(ctx, cells, sigma, tiling, prec):
    s = L.System.quartz(*cells, sigma_ang=sigma, seed=2)
    o = O.System.quartz(*cells, sigma_ang=sigma, seed=2)
    a, b = prepare(ctx, s)
    Mn = 48
    ch = tiling["chunk_cols"]
    last = s.norb - (s.norb - 1) // ch * ch
    assert last < 8 * 64 * tiling["cols_per_lane"]  # the fallback runs
    try:
        ctx.set_tiling(schedule=L.SCHED_BLOCKED, cols_per_lane=1, chunk_cols=4096)
        ref_atoms, _ = ctx.kpe_moments(Mn, L.FP64)
        ctx.set_tiling(**tiling)
        ctx.keep_orbital_moments(True)
        got_atoms, _ = ctx.kpe_moments(Mn, prec)
        j0 = (s.norb - 1) // ch * ch
        cols = np.arange(j0, s.norb, 7, dtype=np.int32)
        got = ctx.orbital_moments(Mn, j0, s.norb)[:, cols - j0].T
    finally:
        ctx.keep_orbital_moments(False)
        ctx.set_tiling()
    ref = o.diag_moments_at(a, b, Mn + 1, cols)
    tol = 1e-12 if prec == L.FP64 else 3e-6
    assert np.abs(got_atoms - ref_atoms).max() < tol
    assert np.abs(got - ref).max() < tol


def test_batch_of_256_cells_at_full_size(ctx):
    """The 256-cell batched workload (86,211 atoms, M = 256, FP32 bulk with a
    per-cell FP64 finish, hipGraph iterations): every cell converges and
    conserves its electrons; eight cells across the size range match their own
    single-structure FP64 solves to 1e-6 e and 1e-6 eV/atom; the smallest cell
    matches the oracle's SCC."""
    from largetb_amd.learned import batch_structures

    cells = batch_structures(256, seed=1)
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in cells]
    assert sum(s.nat for s in systems) == 86211
    with L.Batch(ctx, systems) as b:
        r = b.scc(256, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=400, use_graph=1)
        assert r["status"] == 0 and (r["struct_status"] == 0).all()
        assert r["fp64_iters"] >= 1 and r["graph_launches"] >= 1
        qs = b.split(r["q"])
    for k, s in enumerate(systems):
        assert qs[k].sum() == pytest.approx(s.nelectrons, abs=1e-6)
    sizes = np.array([s.nat for s in systems])
    order = np.argsort(sizes, kind="stable")
    for k in order[np.linspace(0, len(order) - 1, 8).round().astype(int)]:
        ctx.load(systems[k])
        g = ctx.scc(256, precision=L.FP64, tol=1e-6, max_iter=400, use_graph=1)
        assert np.abs(qs[k] - g["q"]).max() < 1e-6, k
        assert abs(r["eband_fp"][k] - g["eband_fp"]) * EV / systems[k].nat < 1e-6, k
    k = int(order[0])
    l, m, n, sg, sd = cells[k]
    ref = O.System.quartz(l, m, n, sigma_ang=sg, seed=sd).scc(256, tol=1e-6, maxit=400)
    assert ref["status"] == 0
    assert np.abs(qs[k] - ref["q"]).max() < 1e-6


def test_largest_cell_density_at_M1024_fp32_against_fp64(ctx):
    """The largest single-GPU cell (16x16x12, 27,648 atoms, 110,592 orbitals,
    sigma 0.08 A) at its configured M = 1024: the neutral-start density in
    both precisions with their autotuned tilings; FP32 charges within 2e-6 e
    and band energy within 1e-6 eV/atom of FP64, electrons conserved."""
    s = L.System.quartz(16, 16, 12, sigma_ang=0.08, seed=1)
    prepare(ctx, s)
    d = {}
    for prec in (L.FP64, L.FP32_ACC64):
        ctx.autotune(prec, 2)
        am, tot = ctx.kpe_moments(1024, prec)
        assert tot[0] == s.norb and np.all(am[0] == 4.0)
        d[prec] = ctx.density(1024)
    dq = 2 * np.abs(d[L.FP32_ACC64]["pop"] - d[L.FP64]["pop"]).max()
    deb = abs(d[L.FP32_ACC64]["eband"] - d[L.FP64]["eband"]) * EV / s.nat
    print(f"27,648 atoms, M=1024: max |dq| {dq:.3e} e, |dE_band| {deb:.3e} eV/atom")
    assert dq < 2e-6 and deb < 1e-6
    for r in d.values():
        assert r["nelec"] == pytest.approx(s.nelectrons, rel=1e-12)
