"""Learned guess: HIP descriptors, forward, backward against the numpy oracle."""
import numpy as np
import pytest

import largetb_amd as L
import oracle as O
from oracle import mlp as OM
from test_oracle_mlp import random_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cells,sigma", [((2, 2, 2), 0.05), ((3, 2, 2), 0.15)])
def test_descriptors_match_oracle(ctx, cells, sigma):
    s = L.System.quartz(*cells, sigma_ang=sigma, seed=3)
    xyz, Z, cell = O.quartz(*cells, sigma_ang=sigma, seed=3)
    f = ctx.descriptors(s)
    ref = OM.descriptors(xyz, Z, cell)
    assert np.abs(f - ref).max() < 2e-5 * max(1.0, np.abs(ref).max())


def dataset(ctx):
    feats, Zs, seg, qtot, tgt = [], [], [0], [], []
    rng = np.random.default_rng(0)
    for cells in [(2, 2, 2), (2, 3, 2), (3, 2, 2)]:
        s = L.System.quartz(*cells, sigma_ang=0.1, seed=sum(cells))
        Z, _ = s.coords()
        feats.append(ctx.descriptors(s))
        Zs.append(Z)
        seg.append(seg[-1] + s.nat)
        qtot.append(s.nelectrons)
        tgt.append(np.where(Z == 14, 0.03, 7.99) + 0.01 * rng.standard_normal(s.nat))
    return (np.concatenate(feats), np.concatenate(Zs), np.array(seg, np.int32), np.array(qtot),
            np.concatenate(tgt))


def test_forward_loss_and_gradient_match_oracle(ctx):
    feat, Z, seg, qtot, t = dataset(ctx)
    m = L.Mlp(ctx, seed=7)
    P = random_params(3, feat.astype(float), Z).astype(np.float32)
    m.set_params(P)
    m.set_data(feat, Z, seg, qtot, t)
    q = m.forward()
    qref = OM.forward(P.astype(float), feat.astype(float), Z, seg, qtot)
    assert np.abs(q - qref).max() < 2e-4
    for s in range(len(seg) - 1):
        assert abs(q[seg[s]:seg[s + 1]].sum() - qtot[s]) < 1e-8
    loss, g = m.loss_grad()
    lref, gref = OM.loss_grad(P.astype(float), feat.astype(float), Z, seg, qtot, t)
    assert loss == pytest.approx(lref, rel=1e-4)
    scale = np.abs(gref).max()
    assert np.abs(g - gref).max() < 2e-4 * scale
    m.close()


def test_gradient_over_many_tiles_matches_oracle(ctx):
    """A 12x12x12 and an 11x12x12 cell, 29,808 atoms: 19,872 O atoms are 311
    tiles of 64, more than the 256 workgroups of the backward pass, so
    workgroups add several tiles; both species end in a ragged tile.  Forward,
    loss and gradient against the oracle."""
    feats, Zs, seg, qtot, tgt = [], [], [0], [], []
    rng = np.random.default_rng(5)
    for cells, seed in (((12, 12, 12), 21), ((11, 12, 12), 22)):
        s = L.System.quartz(*cells, sigma_ang=0.1, seed=seed)
        Z, _ = s.coords()
        feats.append(ctx.descriptors(s))
        Zs.append(Z)
        seg.append(seg[-1] + s.nat)
        qtot.append(s.nelectrons)
        tgt.append(np.where(Z == 14, 2.6, 6.7) + 0.02 * rng.standard_normal(s.nat))
    feat, Z, t = np.concatenate(feats), np.concatenate(Zs), np.concatenate(tgt)
    seg, qtot = np.array(seg, np.int32), np.array(qtot)
    assert (Z == 8).sum() // 64 > 256 and (Z == 8).sum() % 64 != 0 and (Z == 14).sum() % 64 != 0
    m = L.Mlp(ctx, seed=3)
    P = random_params(4, feat.astype(float), Z).astype(np.float32)
    m.set_params(P)
    m.set_data(feat, Z, seg, qtot, t)
    q = m.forward()
    qref = OM.forward(P.astype(float), feat.astype(float), Z, seg, qtot)
    assert np.abs(q - qref).max() < 2e-4
    loss, g = m.loss_grad()
    lref, gref = OM.loss_grad(P.astype(float), feat.astype(float), Z, seg, qtot, t)
    assert loss == pytest.approx(lref, rel=1e-4)
    assert np.abs(g - gref).max() < 2e-4 * np.abs(gref).max()
    # the same data again gives the same gradient bit for bit (fixed-order sums)
    _, g2 = m.loss_grad()
    assert np.array_equal(g, g2)
    m.close()


def test_adam_training_reduces_loss(ctx):
    feat, Z, seg, qtot, t = dataset(ctx)
    m = L.Mlp(ctx, seed=11)
    m.set_data(feat, Z, seg, qtot, t)
    m.fit_normalization()
    hist = m.adam(300, lr=1e-3)
    assert hist[-1] < 0.5 * hist[0]
    s = L.System.quartz(2, 2, 2, sigma_ang=0.1, seed=6)
    p = m.predict(s)
    assert 2 * p.sum() == pytest.approx(s.nelectrons, abs=1e-8)
    m.close()


def test_trained_guess_cuts_iterations_and_keeps_the_answer(ctx):
    """The shipped network on a cell outside its training stream (4x4x3,
    sigma 0.1 A, seed 777): the SCC from the guess converges to the same
    answer as from the neutral start (1e-6 e) in far fewer iterations."""
    from largetb_amd import learned
    s = L.System.quartz(4, 4, 3, sigma_ang=0.1, seed=777)
    ctx.load(s)
    m, info = learned.load(ctx)
    guess = m.predict(s)
    m.close()
    neutral = ctx.scc(256, precision=L.FP64, tol=1e-6, max_iter=300)
    guided = ctx.scc(256, guess=guess, precision=L.FP64, tol=1e-6, max_iter=300)
    assert np.abs(2 * guess - neutral["q"]).max() < 0.02
    assert guided["iters"] < neutral["iters"] - 20
    assert np.abs(guided["q"] - neutral["q"]).max() < 1e-6


def test_context_close_closes_its_batch_and_network(ctx):
    """Closing a context first closes the Batch and Mlp objects made on it
    (they hold its stream); later calls on them raise instead of touching
    freed state.  Short per-atom vectors are refused before the C call."""
    c = L.Context(0)
    s = L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    b = L.Batch(c, [s])
    m = L.Mlp(c)
    c.close()
    assert b._h is None and m._h is None
    with pytest.raises(L.LtbError):
        b.scc(64)
    with pytest.raises(L.LtbError):
        m.params()
    ctx.load(s)
    with pytest.raises(ValueError):
        ctx.scc(64, guess=np.zeros(s.nat - 1))
    with pytest.raises(ValueError):
        ctx.set_populations(np.zeros(s.nat + 1))
