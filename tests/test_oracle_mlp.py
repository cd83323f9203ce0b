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


# ---- pins of the radial block and the periodic image enumeration -------------------
SQRT2PI = np.sqrt(2 * np.pi)


def test_narrow_cell_equals_its_supercell():
    """A 1x1x1 cell (4.25 A between planes) needs images two cells away for
    the 6 A cutoff; its descriptors must equal those of the same atoms in the
    3x3x3 tiling of the cell, which reaches them within one image.  A short
    image reach or a missed periodic shift breaks the equality."""
    xyz, Z, cell = O.quartz(1, 1, 1, sigma_ang=0.1, seed=3)
    cell = np.asarray(cell).reshape(3, 3)
    G = OM.descriptors(xyz, Z, cell)
    tiles = [l * cell[0] + m * cell[1] + n * cell[2] for l in range(3) for m in range(3) for n in range(3)]
    big = np.concatenate([xyz + t for t in tiles])
    Gb = OM.descriptors(big, np.tile(Z, len(tiles)), 3 * cell)
    for k in range(len(tiles)):
        assert np.allclose(Gb[k * len(Z):(k + 1) * len(Z)], G, rtol=1e-11, atol=1e-12)
    # the neighbour counts within 6 A also agree with a search over +-4 images
    pl = OM.pairs(xyz, cell)
    for i in range(len(Z)):
        n = 0
        for l in range(-4, 5):
            for m in range(-4, 5):
                for q in range(-4, 5):
                    d = xyz + l * cell[0] + m * cell[1] + q * cell[2] - xyz[i]
                    r = np.linalg.norm(d, axis=1)
                    n += int(((r < OM.RC) & (r > 1e-12)).sum())
        assert n == len(pl[i])


def _cluster(nat, seed, rmin=1.7 * OM.ANG, rmax=4.8 * OM.ANG):
    """atoms in a 60 Bohr box (no image within the cutoff), all pair
    distances in [rmin, rmax]"""
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < nat:
        p = rng.uniform(-1, 1, 3) * 2.3 * OM.ANG
        if np.linalg.norm(p) > 2.3 * OM.ANG:
            continue
        if all(rmin <= np.linalg.norm(p - q) <= rmax for q in pts):
            pts.append(p)
    xyz = np.array(pts) + 30.0
    Z = np.where(rng.random(nat) < 0.4, 14, 8)
    Z[0], Z[1] = 14, 8
    return xyz, Z, np.eye(3) * 60.0


def test_radial_block_gaussian_sum_rules():
    """Gaussians of width sigma equal to their spacing sum, by the Poisson
    summation formula, to sqrt(2 pi) to within 2 exp(-2 pi^2) = 5e-9 inside
    the grid; the first and second moments over the centres give r and
    r^2 + sigma^2.  So per neighbour species t:
        sum_k G[32t+k]          = sqrt(2 pi) sum_j fc(r_j)
        sum_k mu_k G[32t+k]     = sqrt(2 pi) sum_j r_j fc(r_j)
        sum_k mu_k^2 G[32t+k]   = sqrt(2 pi) sum_j (r_j^2 + sigma^2) fc(r_j)
    with distances from the coordinates directly.  A wrong width, centre
    offset, spacing or species block fails these."""
    xyz, Z, cell = _cluster(7, seed=11)
    G = OM.descriptors(xyz, Z, cell)
    dmu = (OM.RC - OM.R0) / 31.0
    mu = OM.R0 + dmu * np.arange(32)
    for i in range(len(Z)):
        for t, zt in enumerate((14, 8)):
            s0 = s1 = s2 = 0.0
            for j in range(len(Z)):
                if j == i or Z[j] != zt:
                    continue
                r = np.linalg.norm(xyz[j] - xyz[i])
                f = 0.5 * (np.cos(np.pi * r / OM.RC) + 1)
                s0, s1, s2 = s0 + f, s1 + r * f, s2 + (r * r + dmu * dmu) * f
            g = G[i, 32 * t:32 * t + 32]
            assert abs(g.sum() - SQRT2PI * s0) < 1e-7 * max(1.0, s0)
            assert abs((mu * g).sum() - SQRT2PI * s1) < 1e-7 * max(1.0, s1)
            assert abs((mu * mu * g).sum() - SQRT2PI * s2) < 1e-7 * max(1.0, s2)


def test_radial_block_dimer_closed_forms():
    """A Si-O dimer: the cutoff function is 1/2 at rc/2 and 3/4 at rc/3, each
    atom sees only the other species' block, the angular block (two
    neighbours needed) is zero, and nothing is seen at or beyond the cutoff."""
    for r, fc in ((OM.RC / 2, 0.5), (OM.RC / 3, 0.75)):
        xyz = np.array([[30.0, 30.0, 30.0], [30.0 + r, 30.0, 30.0]])
        G = OM.descriptors(xyz, np.array([14, 8]), np.eye(3) * 60.0)
        assert np.all(G[0, :32] == 0) and np.all(G[1, 32:64] == 0)
        assert np.all(G[:, 64:] == 0)
        assert abs(G[0, 32:64].sum() - SQRT2PI * fc) < 1e-8
        assert abs(G[1, 0:32].sum() - SQRT2PI * fc) < 1e-8
        k = int(np.argmax(G[0, 32:64]))
        assert abs(OM.R0 + k * (OM.RC - OM.R0) / 31 - r) <= 0.5 * (OM.RC - OM.R0) / 31 + 1e-12
    for r in (OM.RC, OM.RC * 1.01):
        xyz = np.array([[30.0, 30.0, 30.0], [30.0 + r, 30.0, 30.0]])
        assert np.all(OM.descriptors(xyz, np.array([14, 8]), np.eye(3) * 60.0) == 0)


def test_descriptors_are_invariant_under_translation_rotation_permutation():
    xyz, Z, cell = O.quartz(2, 2, 1, sigma_ang=0.1, seed=9)
    cell = np.asarray(cell).reshape(3, 3)
    G = OM.descriptors(xyz, Z, cell)
    rng = np.random.default_rng(2)
    # translation (atoms may leave the cell: images make that irrelevant)
    Gt = OM.descriptors(xyz + rng.uniform(-5, 5, 3), Z, cell)
    assert np.allclose(Gt, G, rtol=1e-10, atol=1e-12)
    # rotation of atoms and cell together
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    Gr = OM.descriptors(xyz @ Q.T, Z, cell @ Q.T)
    assert np.allclose(Gr, G, rtol=1e-10, atol=1e-12)
    # permutation of the atoms permutes the rows
    perm = rng.permutation(len(Z))
    Gp = OM.descriptors(xyz[perm], Z[perm], cell)
    assert np.allclose(Gp, G[perm], rtol=1e-10, atol=1e-12)
