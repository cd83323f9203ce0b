"""Batched SCC (many independent structures, one launch per Chebyshev step)
against the single-structure SCC and the oracle, through the C ABI.

Per structure the batch runs the same SCC as ltb_scc_run; FP64 results agree
with the single-structure solve to summation order (1e-8 e per charge, which
also covers a one-iteration difference at the 1e-6 e stop: the SCC map
contracts by ~0.0125 per iteration, so one more mixed step moves q_out by
< 1e-8 e)."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O

pytestmark = pytest.mark.gpu
EV = L.HARTREE_EV

# uneven cells (L, M, N, sigma A, seed): 72 to 162 atoms, first one is the
# 72-atom 2x2x2 sigma 0.05 seed 1 workload
CELLS = [(2, 2, 2, 0.05, 1), (3, 2, 2, 0.10, 7), (2, 3, 2, 0.02, 11), (2, 2, 3, 0.15, 3),
         (3, 3, 2, 0.08, 5)]


@pytest.fixture(scope="module")
def systems():
    return [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in CELLS]


@pytest.fixture(scope="module")
def singles(ctx, systems):
    out = []
    for s in systems:
        ctx.load(s)
        out.append(ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300))
    return out


@pytest.fixture(scope="module")
def batch(ctx, systems):
    b = L.Batch(ctx, systems)
    yield b
    b.close()


def test_batch_layout(batch, systems):
    assert batch.S == len(systems)
    assert list(np.diff(batch.offsets)) == [s.nat for s in systems]
    assert batch.nat == sum(s.nat for s in systems)


def test_batch_fp64_matches_single_structures(batch, systems, singles):
    r = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    assert r["status"] == 0 and (r["struct_status"] == 0).all()
    for k, (q, g) in enumerate(zip(batch.split(r["q"]), singles)):
        assert abs(int(r["iters"][k]) - g["iters"]) <= 1, k
        assert np.abs(q - g["q"]).max() < 1e-8, k
        assert r["eband_fp"][k] == pytest.approx(g["eband_fp"], abs=1e-8 * systems[k].nat)
        assert r["mu_ha"][k] == pytest.approx(g["mu_ha"], abs=1e-9)
        assert q.sum() == pytest.approx(systems[k].nelectrons, abs=1e-8)
    # structures converge at different iterations; the batch runs until the last
    assert r["batch_iters"] == r["iters"].max()


def test_batch_first_structure_matches_oracle(batch):
    r = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ref = o.scc(200, tol=1e-6, maxit=300)
    q0 = batch.split(r["q"])[0]
    assert np.abs(q0 - ref["q"]).max() < 1e-8


def test_batch_mixed_precision_matches_fp64(batch, systems, singles):
    """FP32 recursion until a structure's |dq| < 1e-6 e, then FP64 for that
    structure only, converged on an FP64 iteration: charges within 1e-6 e and
    E_band (fixed-point estimate) within 1e-6 eV/atom of the FP64 solve."""
    r = batch.scc(200, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=300)
    assert r["status"] == 0 and (r["struct_status"] == 0).all()
    assert r["fp64_iters"] >= 1
    for k, (q, g) in enumerate(zip(batch.split(r["q"]), singles)):
        assert np.abs(q - g["q"]).max() < 1e-6, k
        assert abs(r["eband_fp"][k] - g["eband_fp"]) * EV / systems[k].nat < 1e-6, k


def test_batch_graph_replays_match_eager(batch):
    for prec, sw in [(L.FP64, 0.0), (L.FP32_ACC64, 1e-4)]:
        eager = batch.scc(200, precision=prec, switch_tol=sw, tol=1e-6, max_iter=300, use_graph=0)
        graph = batch.scc(200, precision=prec, switch_tol=sw, tol=1e-6, max_iter=300, use_graph=1)
        assert graph["graph_launches"] >= 1
        assert np.array_equal(graph["iters"], eager["iters"])
        assert np.array_equal(graph["q"], eager["q"])
        assert np.array_equal(graph["eband_fp"], eager["eband_fp"])


def test_batch_converged_guess_stops_after_one_iteration(batch):
    r = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    again = batch.scc(200, guess=r["pop"], precision=L.FP64, tol=1e-6, max_iter=300)
    assert (again["iters"] == 1).all()
    assert np.abs(again["q"] - r["q"]).max() < 1e-6


def test_batch_iteration_cap(batch):
    r = batch.scc(200, precision=L.FP64, max_iter=3, check=False)
    assert r["status"] == L.LTB_NOT_CONVERGED
    assert (r["struct_status"] == L.LTB_NOT_CONVERGED).all() and (r["iters"] == 3).all()


def test_batch_nonfinite_structure_is_isolated(batch, singles):
    """A NaN in one structure's guess fails that structure only."""
    p = np.concatenate([g["pop"] for g in singles])
    bad = p.copy()
    bad[batch.offsets[2] + 3] = np.nan
    r = batch.scc(200, guess=bad, precision=L.FP64, tol=1e-6, max_iter=300, check=False)
    assert r["status"] == L.LTB_ERR_NONFINITE
    st = r["struct_status"]
    assert st[2] == L.LTB_ERR_NONFINITE and (np.delete(st, 2) == 0).all()
    for k in (0, 1, 3, 4):
        assert np.abs(batch.split(r["q"])[k] - singles[k]["q"]).max() < 1e-6


def test_batch_of_uneven_training_cells(ctx):
    """32 cells drawn as the training set (2x2x2 .. 5x5x4, sigma 0.02-0.15 A),
    M=256, FP32 bulk with FP64 finish: every cell converges, conserves its
    electrons, and three of them match their own single-structure FP64 solve."""
    from largetb_amd.learned import batch_structures

    cells = batch_structures(32, seed=99)
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in cells]
    with L.Batch(ctx, systems) as b:
        r = b.scc(256, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=400, use_graph=1)
        assert r["status"] == 0 and (r["struct_status"] == 0).all()
        qs = b.split(r["q"])
    for k, s in enumerate(systems):
        assert qs[k].sum() == pytest.approx(s.nelectrons, abs=1e-6)
    for k in (0, 13, 31):
        ctx.load(systems[k])
        g = ctx.scc(256, precision=L.FP64, tol=1e-6, max_iter=400, use_graph=1)
        assert np.abs(qs[k] - g["q"]).max() < 1e-6, k


def test_batch_tilings_agree_and_autotune(batch):
    """Every tiling of the batched step kernel gives the same FP64 SCC (the
    grouping of the FP64 partial sums differs, so 1e-10 e), and the autotuner
    leaves a valid tiling behind."""
    base = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    for apw, bpc, nt in [(4, 2, 0), (16, 8, 1), (1, 1, 1)]:
        batch.set_tiling(atoms_per_wave=apw, blocks_per_cu=bpc, nontemporal=nt)
        t = batch.tiling()
        assert (t["atoms_per_wave"], t["blocks_per_cu"], t["nontemporal"]) == (apw, bpc, nt)
        r = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
        assert np.array_equal(r["iters"], base["iters"])
        assert np.abs(r["q"] - base["q"]).max() < 1e-10
    best, gbs = batch.autotune(L.FP32_ACC64, 2)
    assert gbs > 0 and best["atoms_per_wave"] in (4, 8, 16) and best["blocks_per_cu"] in (2, 4, 8)
    r = batch.scc(200, precision=L.FP64, tol=1e-6, max_iter=300)
    assert np.abs(r["q"] - base["q"]).max() < 1e-10


def test_batch_covers_every_column_group_shape(ctx):
    """Cells whose orbital counts leave 0, 64, 128 and 256 columns for the tail
    group (4x4x4: 2304 = 9 x 256; 2x2x2: 288; 2x3x4: 864; 3x3x2: 648), in both
    precisions, against their own single-structure solves."""
    cells = [(4, 4, 4, 0.05, 21), (2, 2, 2, 0.08, 22), (2, 3, 4, 0.05, 23), (3, 3, 2, 0.12, 24)]
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in cells]
    assert sorted((s.norb % 256) for s in systems) == [0, 32, 96, 136]
    singles = []
    for s in systems:
        ctx.load(s)
        singles.append(ctx.scc(128, precision=L.FP64, tol=1e-6, max_iter=300))
    with L.Batch(ctx, systems) as b:
        r64 = b.scc(128, precision=L.FP64, tol=1e-6, max_iter=300)
        rmx = b.scc(128, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=300)
        q64, qmx = b.split(r64["q"]), b.split(rmx["q"])
    assert r64["status"] == 0 and rmx["status"] == 0
    for k, g in enumerate(singles):
        assert np.abs(q64[k] - g["q"]).max() < 1e-8, k
        assert np.abs(qmx[k] - g["q"]).max() < 1e-6, k
