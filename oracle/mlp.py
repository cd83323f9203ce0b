"""Numpy float64 restatement of the learned-guess descriptor and network.

TEST INFRASTRUCTURE ONLY (the checker of the HIP kernels in
largetb_amd/csrc/mlp.hip).  The model has no counterpart in the reference;
its definition is the project's:
  descriptor: 2 x 32 radial Gaussians per neighbour species and 16 angular
  terms, cutoff 6 A; network per centre species 80-64-64-64-1 with SiLU;
  q_raw = shift_c + y; mean projection q = q_raw - (sum q_raw - Q)/n.
Pins (tests/test_oracle_mlp.py): backward against central finite
differences of forward; the angular term against a direct triple loop on a
small cell; the radial block by the Poisson sum rules of unit-spacing
Gaussians (sum, first and second moment over the centres) on a cluster and
by dimer closed forms (fc = 1/2 at rc/2, 3/4 at rc/3, species blocks, zero
at the cutoff); the image enumeration by a 1x1x1 cell (images two cells
away) equalling its 3x3x3 tiling and a +-4 image count; translation,
rotation and permutation invariance; the projection against its definition
(sum q = Q).
"""
import numpy as np

ANG = 1.88973
RC = 6.0 * ANG
R0 = 0.5 * ANG
NF, NH = 80, 64
OW1, OB1 = 0, NH * NF
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH * NH
OW4 = OB3 + NH
OB4 = OW4 + NH
PS = OB4 + 1
NTRAIN = 2 * PS
OMEAN = NTRAIN
OSD = OMEAN + 2 * NF
OSHIFT = OSD + 2 * NF
NPARAM = OSHIFT + 2
ZETA = [1, 2, 4, 8]
LAMBDA = [-1.0, 1.0]
ETAA = [0.01, 0.04]


def pairs(xyz, cell, rc=RC):
    """every (i, j, image) with 0 < r < rc; images as far as rc reaches"""
    xyz = np.asarray(xyz, float)
    cell = np.asarray(cell, float).reshape(3, 3)
    n = len(xyz)
    vol = abs(np.linalg.det(cell))
    widths = [vol / np.linalg.norm(np.cross(cell[(d + 1) % 3], cell[(d + 2) % 3])) for d in range(3)]
    reach = [int(np.ceil(rc / w)) + 1 for w in widths]
    out = [[] for _ in range(n)]
    rng = [range(-r, r + 1) for r in reach]
    shifts = np.array([l * cell[0] + m * cell[1] + q * cell[2] for l in rng[0] for m in rng[1] for q in rng[2]])
    for i in range(n):
        d = xyz[None, :, :] + shifts[:, None, :] - xyz[i]        # [img, j, 3]
        r = np.linalg.norm(d, axis=2)
        img, j = np.nonzero((r < rc) & (r > 1e-12))
        for a, b in zip(img, j):
            out[i].append((b, d[a, b]))
    return out


def descriptors(xyz, Z, cell):
    Z = np.asarray(Z)
    sp = np.where(Z == 14, 0, 1)
    pl = pairs(xyz, cell)
    dmu = (RC - R0) / 31.0
    eta = 1.0 / (2 * dmu * dmu)
    mu = R0 + dmu * np.arange(32)
    G = np.zeros((len(Z), NF))
    for i, lst in enumerate(pl):
        if not lst:
            continue
        js = np.array([p[0] for p in lst])
        D = np.array([p[1] for p in lst])
        r = np.linalg.norm(D, axis=1)
        fc = 0.5 * (np.cos(np.pi * r / RC) + 1)
        rad = np.exp(-eta * (r[:, None] - mu[None, :]) ** 2) * fc[:, None]
        for t in (0, 1):
            G[i, 32 * t:32 * t + 32] = rad[sp[js] == t].sum(0)
        cos = (D @ D.T) / np.outer(r, r)
        r2 = r[:, None] ** 2 + r[None, :] ** 2
        ff = np.outer(fc, fc)
        iu = np.triu_indices(len(r), 1)
        for z, zeta in enumerate(ZETA):
            for l, lam in enumerate(LAMBDA):
                for e, ea in enumerate(ETAA):
                    term = 2.0 ** (1 - zeta) * (1 + lam * cos) ** zeta * np.exp(-ea * r2) * ff
                    G[i, 64 + 4 * z + 2 * l + e] = term[iu].sum()
    return G


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def _net(P, c):
    W = P[c * PS:(c + 1) * PS]
    return (W[OW1:OB1].reshape(NH, NF), W[OB1:OW2], W[OW2:OB2].reshape(NH, NH), W[OB2:OW3],
            W[OW3:OB3].reshape(NH, NH), W[OB3:OW4], W[OW4:OB4], W[OB4])


def forward(P, feat, Z, seg, qtot, keep=False):
    P = np.asarray(P, float)
    Z = np.asarray(Z)
    n = len(Z)
    qraw = np.zeros(n)
    cache = {}
    for c in (0, 1):
        idx = np.nonzero(np.where(Z == 14, 0, 1) == c)[0]
        if not len(idx):
            continue
        W1, b1, W2, b2, W3, b3, w4, b4 = _net(P, c)
        x = (feat[idx] - P[OMEAN + c * NF:OMEAN + (c + 1) * NF]) / P[OSD + c * NF:OSD + (c + 1) * NF]
        z1 = x @ W1.T + b1
        z2 = silu(z1) @ W2.T + b2
        z3 = silu(z2) @ W3.T + b3
        y = silu(z3) @ w4 + b4
        qraw[idx] = P[OSHIFT + c] + y
        cache[c] = (idx, x, z1, z2, z3)
    q = qraw.copy()
    for s in range(len(seg) - 1):
        a, b = seg[s], seg[s + 1]
        q[a:b] -= (qraw[a:b].sum() - qtot[s]) / (b - a)
    return (q, cache) if keep else q


def loss_grad(P, feat, Z, seg, qtot, target):
    P = np.asarray(P, float)
    q, cache = forward(P, feat, Z, seg, qtot, keep=True)
    n = len(Z)
    loss = ((q - target) ** 2).sum() / n
    g = 2 * (q - target) / n
    for s in range(len(seg) - 1):
        a, b = seg[s], seg[s + 1]
        g[a:b] -= g[a:b].mean()
    grad = np.zeros(NTRAIN)
    for c, (idx, x, z1, z2, z3) in cache.items():
        W1, b1, W2, b2, W3, b3, w4, b4 = _net(P, c)
        dy = g[idx]
        G = grad[c * PS:(c + 1) * PS]
        G[OW4:OB4] = silu(z3).T @ dy
        G[OB4] = dy.sum()
        d3 = dy[:, None] * w4[None, :] * dsilu(z3)
        G[OW3:OB3] = (d3.T @ silu(z2)).ravel()
        G[OB3:OW4] = d3.sum(0)
        d2 = (d3 @ W3) * dsilu(z2)
        G[OW2:OB2] = (d2.T @ silu(z1)).ravel()
        G[OB2:OW3] = d2.sum(0)
        d1 = (d2 @ W2) * dsilu(z1)
        G[OW1:OB1] = (d1.T @ x).ravel()
        G[OB1:OW2] = d1.sum(0)
    return loss, grad
