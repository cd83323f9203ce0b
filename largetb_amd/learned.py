"""Learned SCC starting guess: training-set generation, training and use.

The training structures follow the batched workload: alpha-quartz supercells
with L, M in {2..5}, N in {2..4} and a displacement sigma uniform in
[0.02, 0.15] A, drawn from a SplitMix64 stream.  Each is solved from the
neutral start without any guess: either one at a time by the FP64 SCC
(`generate`) or 256 at a time by the batched SCC (`solve_batched`: the
two-level SCC with low-order iterations at M_lo = 32, an FP32 bulk and an FP64
finish; charges within ~3e-8 e of each structure's FP64 solve).  The
descriptors and converged charges form the data.  Training runs the device
kernels (ltb_mlp_adam).
"""
import json
import os
import time

import numpy as np

from . import FP32_ACC64, FP64, Batch, Context, Mlp, System

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


def solve_batched(ctx, structures, M=256, kT_ev=0.05, per_batch=256, log=None, M_lo=32):
    """Converged charges of every structure, `per_batch` structures per batched
    SCC (FP32 bulk, FP64 finish, |dq| < 1e-6 e; with M_lo > 0 the bulk is the
    two-level SCC, low-order iterations at M_lo).  Returns the per-atom charges
    (e, structure order) and per-structure atom offsets, iterations, status."""
    qs, offs, iters, status = [], [0], [], []
    t0 = time.perf_counter()
    for b0 in range(0, len(structures), per_batch):
        systems = [System.quartz(L, Mc, N, sigma_ang=sg, seed=sd)
                   for L, Mc, N, sg, sd in structures[b0:b0 + per_batch]]
        with Batch(ctx, systems) as b:
            if b0 == 0:
                b.autotune(FP32_ACC64, 3)
            r = b.scc(M, kT_ev=kT_ev, precision=FP32_ACC64, switch_tol=1e-6, tol=1e-6,
                      max_iter=2000 if M_lo else 400, use_graph=1, check=False, M_lo=M_lo)
        qs.append(r["q"])
        for s in systems:
            offs.append(offs[-1] + s.nat)
        iters.extend(r["iters"].tolist())
        status.extend(r["struct_status"].tolist())
        if log:
            log(f"batched solve: {min(b0 + per_batch, len(structures))}/{len(structures)} structures, "
                f"{offs[-1]} atoms, {time.perf_counter() - t0:.1f} s")
    return (np.concatenate(qs), np.array(offs, np.int64), np.array(iters, np.int32),
            np.array(status, np.int32))


def assemble(ctx, structures, q, offs, status, log=None):
    """Descriptors and targets of the converged structures (status 0)."""
    feats, Zs, qk, seg, qtot = [], [], [], [0], []
    t0 = time.perf_counter()
    for k, (L, Mc, N, sg, sd) in enumerate(structures):
        if log and k % 2048 == 0:
            log(f"descriptors: {k}/{len(structures)} structures, {time.perf_counter() - t0:.1f} s")
        if status[k] != 0:
            continue
        s = System.quartz(L, Mc, N, sigma_ang=sg, seed=sd)
        Z, _ = s.coords()
        feats.append(ctx.descriptors(s))
        Zs.append(Z)
        qk.append(q[offs[k]:offs[k + 1]])
        seg.append(seg[-1] + s.nat)
        qtot.append(s.nelectrons)
    return dict(feat=np.concatenate(feats), Z=np.concatenate(Zs), q=np.concatenate(qk),
                seg=np.array(seg, np.int32), qtot=np.array(qtot))


def train(ctx, data, steps=3000, lr=2e-3, seed=1, log=None, seconds=None, init=None):
    """Full-batch Adam on the device; lr drops 4x at half the steps (or half
    of `seconds`).  With `seconds`, training stops at the first 100-step
    boundary past that time.  `init`: parameters to continue from."""
    m = Mlp(ctx, seed=seed)
    m.set_data(data["feat"], data["Z"], data["seg"], data["qtot"], data["q"])
    if init is None:
        m.fit_normalization()
    else:  # continue from trained parameters (their feature normalisation included)
        m.set_params(init)
    hist = []
    done = 0
    t0 = time.perf_counter()
    while done < steps and (seconds is None or time.perf_counter() - t0 < seconds):
        k = min(100, steps - done)
        late = done >= steps // 2 or (seconds is not None and time.perf_counter() - t0 > seconds / 2)
        h = m.adam(k, lr=lr / 4 if late else lr)
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
    generates the training set on the GPU, trains, and writes guess_weights.npz.
    The shipped guess_weights.npz was made with --batched --M 256 over 20,480
    structures (two generation runs of 10,240 with --charges-out, then
    --charges-in both --init guess_weights_256.npz --steps 1700);
    guess_weights_256.npz is the default recipe's 256-structure network.
    --batched solves 256 structures per batched SCC; --charges-out writes only
    the converged charges (with --skip-train, a generation-only run);
    --charges-in trains from charges files written that way (several files =
    consecutive slices of the same structure stream)."""
    import argparse
    import sys
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=256)
    ap.add_argument("--first", type=int, default=0, help="skip this many structures of the stream")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--M", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--train-seconds", type=float, default=None)
    ap.add_argument("--init", default="", help="continue training from these weights (.npz)")
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--charges-out", default="")
    ap.add_argument("--charges-in", nargs="*", default=[])
    ap.add_argument("--skip-train", action="store_true")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--data-out", default="")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    ctx = Context(0)
    t0 = time.perf_counter()
    if a.charges_in:
        parts = [np.load(p, allow_pickle=False) for p in a.charges_in]
        first = int(parts[0]["first"])
        count = sum(int(p["count"]) for p in parts)
        structures = batch_structures(first + count, a.seed)[first:]
        q = np.concatenate([p["q"].astype(np.float64) for p in parts])
        offs = [0]
        for p in parts:
            offs.extend((offs[-1] + p["offs"][1:]).tolist())
        status = np.concatenate([p["status"] for p in parts])
        iters = np.concatenate([p["iters"] for p in parts])
        data = assemble(ctx, structures, q, np.array(offs), status, log=log)
        M = int(parts[0]["M"])
    elif a.batched:
        structures = batch_structures(a.first + a.count, a.seed)[a.first:]
        q, offs, iters, status = solve_batched(ctx, structures, M=a.M, log=log)
        if a.charges_out:
            np.savez_compressed(a.charges_out, q=q.astype(np.float32), offs=offs, iters=iters,
                                status=status, first=a.first, count=a.count, seed=a.seed, M=a.M)
        if a.skip_train:
            print(json.dumps(dict(count=a.count, first=a.first, atoms=int(offs[-1]),
                                  converged=int((status == 0).sum()), mean_iters=float(iters.mean()),
                                  seconds=round(time.perf_counter() - t0, 1))))
            ctx.close()
            return
        data = assemble(ctx, structures, q, offs, status, log=log)
        M = a.M
    else:
        data = generate(ctx, batch_structures(a.count, a.seed), M=a.M, log=log)
        iters = np.array([x["iters"] for x in data["meta"]])
        M = a.M
    tgen = time.perf_counter() - t0
    log(f"data: {len(data['seg']) - 1} structures, {len(data['Z'])} atoms, {tgen:.1f} s")
    if a.data_out:
        np.savez_compressed(a.data_out, feat=data["feat"], Z=data["Z"], q=data["q"], seg=data["seg"],
                            qtot=data["qtot"])
    t0 = time.perf_counter()
    init = np.load(a.init, allow_pickle=False)["params"] if a.init else None
    m, hist = train(ctx, data, steps=a.steps, lr=a.lr, log=log, seconds=a.train_seconds, init=init)
    ttrain = time.perf_counter() - t0
    q = m.forward()
    err = np.abs(q - data["q"])
    info = dict(count=len(data["seg"]) - 1, seed=a.seed, M=M, atoms=int(len(data["Z"])), steps=len(hist),
                train_mse=float(hist[-1]), train_max_err=float(err.max()),
                generate_s=round(tgen, 1), train_s=round(ttrain, 1),
                mean_scc_iters=float(np.mean(iters)))
    save(m, a.out, info)
    print(json.dumps(info))
    m.close()
    ctx.close()


if __name__ == "__main__":
    main()
