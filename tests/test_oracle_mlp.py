"""Pins of the numpy learned-guess oracle (CPU only)."""
import numpy as np

import oracle as O
from oracle import mlp as OM


def small_set(seed=0):
    rng = np.random.default_rng(seed)
    feats, Zs, seg, qtot, tgt = [], [], [0], [], []
    for cells in [(2, 2, 2), (2, 3, 2)]:
        xyz, Z, cell = O.quartz(*cells, sigma_ang=0.1, seed=seed + sum(cells))
        feats.append(OM.descriptors(xyz, Z, cell))
        Zs.append(Z)
        seg.append(seg[-1] + len(Z))
        qtot.append(float(np.where(Z == 14, 4.0, 6.0).sum()))
        tgt.append(np.where(Z == 14, 0.03, 7.99) + 0.01 * rng.standard_normal(len(Z)))
    return np.concatenate(feats), np.concatenate(Zs), np.array(seg), np.array(qtot), np.concatenate(tgt)


def random_params(seed=1, feat=None, Z=None):
    rng = np.random.default_rng(seed)
    P = np.zeros(OM.NPARAM)
    P[:OM.NTRAIN] = 0.2 * rng.standard_normal(OM.NTRAIN)
    P[OM.OSD:OM.OSD + 2 * OM.NF] = 1.0
    if feat is not None:
        for c, z in enumerate((14, 8)):
            f = feat[Z == z]
            P[OM.OMEAN + c * OM.NF:OM.OMEAN + (c + 1) * OM.NF] = f.mean(0)
            P[OM.OSD + c * OM.NF:OM.OSD + (c + 1) * OM.NF] = np.maximum(f.std(0), 1e-3)
    P[OM.OSHIFT:OM.OSHIFT + 2] = [0.03, 7.99]
    return P


def test_backward_matches_finite_differences():
    feat, Z, seg, qtot, t = small_set()
    P = random_params(1, feat, Z)
    loss, g = OM.loss_grad(P, feat, Z, seg, qtot, t)
    rng = np.random.default_rng(5)
    for j in rng.choice(OM.NTRAIN, 40, replace=False):
        h = 1e-6
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += h
        Pm[j] -= h
        lp, _ = OM.loss_grad(Pp, feat, Z, seg, qtot, t)
        lm, _ = OM.loss_grad(Pm, feat, Z, seg, qtot, t)
        assert abs((lp - lm) / (2 * h) - g[j]) < 1e-7 + 1e-5 * abs(g[j])


def test_projection_is_neutral_per_structure():
    feat, Z, seg, qtot, t = small_set()
    q = OM.forward(random_params(2, feat, Z), feat, Z, seg, qtot)
    for s in range(len(seg) - 1):
        assert abs(q[seg[s]:seg[s + 1]].sum() - qtot[s]) < 1e-9


def test_angular_term_against_direct_triple_loop():
    xyz, Z, cell = O.quartz(2, 2, 2, sigma_ang=0.1, seed=4)
    G = OM.descriptors(xyz, Z, cell)
    pl = OM.pairs(xyz, cell)
    for i in (0, 5, 40):
        D = [p[1] for p in pl[i]]
        ref = np.zeros(16)
        for a in range(len(D)):
            for b in range(a + 1, len(D)):
                ra, rb = np.linalg.norm(D[a]), np.linalg.norm(D[b])
                ct = D[a] @ D[b] / (ra * rb)
                fa = 0.5 * (np.cos(np.pi * ra / OM.RC) + 1)
                fb = 0.5 * (np.cos(np.pi * rb / OM.RC) + 1)
                for z, zeta in enumerate(OM.ZETA):
                    for l, lam in enumerate(OM.LAMBDA):
                        for e, ea in enumerate(OM.ETAA):
                            ref[4 * z + 2 * l + e] += 2.0 ** (1 - zeta) * (1 + lam * ct) ** zeta * \
                                np.exp(-ea * (ra * ra + rb * rb)) * fa * fb
        assert np.allclose(G[i, 64:], ref, rtol=1e-12, atol=1e-12)
    # neighbour count within 6 A of an O in quartz is ~50-60 (including images)
    assert 30 < len(pl[5]) < 90
