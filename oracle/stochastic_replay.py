"""Replay of the reference's stochastic estimator from exact eigenvalues.

TEST INFRASTRUCTURE ONLY (numpy).  It restates, for one structure with known
eigenvalues e_k of H, the chain the reference runs in stochastic mode:
  runKPM (withSP2_deltaH.cpp:1528-1723): Gershgorin frame, 1000 trace moments
    over nRandoms Gaussian vectors, Jackson DOS at 2000 Chebyshev nodes, mu at
    the Ne/(2 N) fraction of the DOS area, bounds where DOS/N >= 0.001, the
    narrowed frame;
  getKPE's step coefficients at mu with the gmMs kernel (:1094-1132, M = 200);
  computeAlphas' trace (:1240-1281): sum_r r^T P r / nRandoms.
For Gaussian r and an orthonormal eigenbasis V, c = V^T r has iid N(0, 1)
entries, so r^T f(H) r = sum_k c_k^2 f(e_k) and the sum over nRandoms vectors
weights eigenstate k by w_k ~ chi^2(nRandoms), independently over k.  A draw
of the reference's estimator therefore needs only the eigenvalues: the
replay is exact in distribution (the reference's own draws come from
std::random_device and cannot be repeated).
"""
import json

import numpy as np

import oracle as O


def cheb_moments(x, w, nm):
    """sum_k w_k T_m(x_k), m < nm"""
    out = np.zeros(nm)
    t0, t1 = np.ones_like(x), x.copy()
    out[0], out[1] = w.sum(), (w * t1).sum()
    for m in range(2, nm):
        t0, t1 = t1, 2 * x * t1 - t0
        out[m] = (w * t1).sum()
    return out


def step_occupation(x, muval, M=200):
    """getKPE's P as a function of the scaled eigenvalue: c_0/2 + sum c_m g_m T_m"""
    coef = O.kpe_weights(O.step_coeffs(M, muval), O.gmM(M))
    t0, t1 = np.ones_like(x), x.copy()
    s = coef[0] * t0 + coef[1] * t1
    for k in range(2, M):
        t0, t1 = t1, 2 * x * t1 - t0
        s += coef[k] * t1
    return s


def replay(ev, gersh, nelectrons, nrandoms=500, nmoments=1000, M=200, ndraw=21, seed=12345, narrow=True):
    """ndraw independent runKPM + computeAlphas draws; returns the traces the
    reference would print ("sp2count is 0, and trace is ..."), the mu values
    and the bounds of each draw"""
    rng = np.random.default_rng(seed)
    n = len(ev)
    a0, b0 = (gersh[1] - gersh[0]) / 2, (gersh[1] + gersh[0]) / 2
    out = []
    for _ in range(ndraw):
        w = rng.chisquare(nrandoms, n)
        d = O.kpm_dos(cheb_moments((ev - b0) / a0, w, nmoments), a0, b0, n, nelectrons)
        a, b = ((d["emax"] - d["emin"]) / 2, (d["emax"] + d["emin"]) / 2) if narrow else (a0, b0)
        f = step_occupation((ev - b) / a, (d["mu"] - b) / a, M)
        w2 = rng.chisquare(nrandoms, n)
        out.append(dict(trace=float((w2 * f).sum() / nrandoms), trace_exact=float(f.sum()), mu=float(d["mu"]),
                        emin=float(d["emin"]), emax=float(d["emax"])))
    return out


def main(argv=None):
    """python -m oracle.stochastic_replay L [--zper]: the reference-lattice LxLxL
    cell (x/y images as the reference, or all three with --zper), exact
    eigenvalues by numpy, 21 draws with nRandoms = 500 (./scf L L L 500)"""
    import sys
    argv = sys.argv[1:] if argv is None else argv
    L = int(argv[0]) if argv else 8
    if "--zper" in argv:
        xyz, Z, cell = O.quartz(L, L, L, mode=O.MODE_REFERENCE)
        o = O.System(xyz, Z, cell, periodic_mask=7)
    else:
        o = O.System.quartz(L, L, L, mode=O.MODE_REFERENCE)
    ev = np.linalg.eigvalsh(o.dense_h())
    gersh = o.gershgorin()
    ne = 48.0 * L ** 3  # main :1766
    res = replay(ev, gersh, ne)
    tr = [r["trace"] for r in res]
    print(json.dumps({"cell": f"{L}x{L}x{L} reference lattice{' z-periodic' if '--zper' in argv else ''}",
                      "orbitals": len(ev), "Ne/2": ne / 2, "gershgorin": list(gersh),
                      "spectrum": [float(ev[0]), float(ev[-1])],
                      "homo_lumo": [float(ev[int(ne / 2) - 1]), float(ev[int(ne / 2)])],
                      "bounds_after_narrowing": sorted({(round(r["emin"], 6), round(r["emax"], 6)) for r in res}),
                      "mu_values": sorted({round(r["mu"], 6) for r in res}),
                      "trace_min": min(tr), "trace_max": max(tr), "trace_mean": float(np.mean(tr)),
                      "trace_sd": float(np.std(tr)), "traces": [round(t, 1) for t in tr]}))


if __name__ == "__main__":
    main()
