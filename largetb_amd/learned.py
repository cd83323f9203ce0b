"""Learned SCC starting guess: training-set generation, training and use.

The training structures follow the batched workload: alpha-quartz supercells
with L, M in {2..5}, N in {2..4} and a displacement sigma uniform in
[0.02, 0.15] A, drawn from a SplitMix64 stream.  Each is solved by the plain
SCC (FP64, neutral start); its descriptors and converged charges form the
data.  Training runs the device kernels (ltb_mlp_adam).
"""
import json
import os
import time

import numpy as np

from . import FP64, Context, Mlp, System

MASK = (1 << 64) - 1
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "guess_weights.npz")


def _splitmix(state):
    state = (state + 0x9E3779B97F4A7C15) & MASK
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return state, z ^ (z >> 31)


def batch_structures(count, seed):
    """(L, M, N, sigma_ang, seed) of `count` uneven quartz cells."""
    st = seed & MASK
    out = []
    for _ in range(count):
        st, a = _splitmix(st)
        st, b = _splitmix(st)
        st, c = _splitmix(st)
        st, d = _splitmix(st)
        st, e = _splitmix(st)
        sigma = 0.02 + 0.13 * ((d >> 11) * 2.0 ** -53)
        out.append((2 + a % 4, 2 + b % 4, 2 + c % 3, sigma, e))
    return out


def generate(ctx, structures, M=512, kT_ev=0.05, log=None):
    """Solve every structure with the plain FP64 SCC; return the data set."""
    feats, Zs, qs, seg, qtot, meta = [], [], [], [0], [], []
    t0 = time.perf_counter()
    for k, (L, Mc, N, sigma, seed) in enumerate(structures):
        s = System.quartz(L, Mc, N, sigma_ang=sigma, seed=seed)
        ctx.load(s)
        r = ctx.scc(M, kT_ev=kT_ev, precision=FP64, tol=1e-6, max_iter=400, use_graph=1, check=False)
        if r["status"] != 0:
            if log:
                log(f"structure {k} {L}x{Mc}x{N} did not converge (status {r['status']}); skipped")
            continue
        Z, _ = s.coords()
        feats.append(ctx.descriptors(s))
        Zs.append(Z)
        qs.append(r["q"])
        seg.append(seg[-1] + s.nat)
        qtot.append(s.nelectrons)
        meta.append(dict(cells=[L, Mc, N], sigma=sigma, iters=r["iters"]))
        if log and (k % 16 == 0 or k == len(structures) - 1):
            log(f"training set: {k + 1}/{len(structures)} structures, {seg[-1]} atoms, "
                f"{time.perf_counter() - t0:.1f} s")
    return dict(feat=np.concatenate(feats), Z=np.concatenate(Zs), q=np.concatenate(qs),
                seg=np.array(seg, np.int32), qtot=np.array(qtot), meta=meta)


def train(ctx, data, steps=3000, lr=2e-3, seed=1, log=None):
    m = Mlp(ctx, seed=seed)
    m.set_data(data["feat"], data["Z"], data["seg"], data["qtot"], data["q"])
    m.fit_normalization()
    hist = []
    done = 0
    while done < steps:
        k = min(500, steps - done)
        h = m.adam(k, lr=lr if done < steps // 2 else lr / 4)
        hist.extend(h.tolist())
        done += k
        if log:
            log(f"adam {done}/{steps}: mse {h[-1]:.3e} e^2")
    return m, np.array(hist)


def save(m, path=WEIGHTS, info=None):
    np.savez_compressed(path, params=m.params(), info=json.dumps(info or {}))


def load(ctx, path=WEIGHTS):
    d = np.load(path, allow_pickle=False)
    m = Mlp(ctx)
    m.set_params(d["params"])
    return m, json.loads(str(d["info"]))


def main(argv=None):
    """python -m largetb_amd.learned [--count 256] [--seed 2026] [--M 512] [--steps 3000]
    generates the training set on the GPU, trains, and writes guess_weights.npz"""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--M", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--data-out", default="")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = Context(0)
    t0 = time.perf_counter()
    data = generate(ctx, batch_structures(a.count, a.seed), M=a.M, log=log)
    tgen = time.perf_counter() - t0
    if a.data_out:
        np.savez_compressed(a.data_out, feat=data["feat"], Z=data["Z"], q=data["q"], seg=data["seg"],
                            qtot=data["qtot"])
    t0 = time.perf_counter()
    m, hist = train(ctx, data, steps=a.steps, log=log)
    ttrain = time.perf_counter() - t0
    q = m.forward()
    err = np.abs(q - data["q"])
    info = dict(count=a.count, seed=a.seed, M=a.M, atoms=int(len(data["Z"])), steps=a.steps,
                train_mse=float(hist[-1]), train_max_err=float(err.max()),
                generate_s=round(tgen, 1), train_s=round(ttrain, 1),
                mean_scc_iters=float(np.mean([x["iters"] for x in data["meta"]])))
    save(m, a.out, info)
    print(json.dumps(info))
    m.close()
    ctx.close()


if __name__ == "__main__":
    main()
