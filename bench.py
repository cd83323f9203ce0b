"""Benchmark of the KPE / SCC density-matrix path on one MI355X.

Workload (default): alpha-quartz 12x12x12 supercell, 15,552 atoms, 62,208
orbitals, sigma 0.10 A, exact-trace KPE with M=512, Fermi-Dirac kT=0.05 eV,
mu bisection, populations, Coulomb shift and linear mixing 0.2.  One step =
one SCC iteration (FP32 Chebyshev recursion with FP64 moment accumulation,
everything else FP64), inputs resident in HBM.

value = algorithmic Chebyshev block-SpMM bytes of the timed iterations divided
by their wall time (GB/s).  roofline = the block-SpMM kernel's algorithmic
bytes per launch over its mean launch time (HIP events on the context stream
around every launch of the first timed iteration).
time_to_solution (default on, --no-tts): plain FP64 SCC from the neutral
start against the learned guess with the two-level SCC (low-order iterations
at M=32 corrected by the defect of the last full-order iteration, FP32
full-order iterations until |dq| < 1e-2 e, FP64 after), then, as the time
budget allows, against the learned guess with FP32 bulk and FP64 finish; all
to |dq| < 1e-6 e on a full-order FP64 iteration, within --time-budget seconds
of the start (what would not fit is skipped, saying so); one line per SCC
iteration goes to stderr meanwhile.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "SCC time-to-solution, 15.5k-atom SiO2: learned+mixed vs plain FP64; SpMM GB/s"
TWO_LEVEL_M_LO = 32  # order of the low-order iterations of the two-level SCC
HBM_PEAK_GBS = 8000.0
T_START = time.perf_counter()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Threads for the CPU leg: the process's CPU share as OMP_NUM_THREADS
    states it, else every core the process may run on."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    if n > 0:
        return n
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(cells, sigma, seed, M, gpu_bytes_per_iter, target_s=12.0):
    """The oracle's op-for-op FP64 restatement of getKPE (CSR SpMV + BLAS-1 per
    step) on a sample of unit-vector columns of the same structure.

    value: the GPU line's algorithmic bytes of one SCC iteration divided by the
    CPU's seconds per SCC iteration (the sample's time per column times all
    columns), i.e. the same work unit as the GPU value."""
    import numpy as np
    import oracle as O
    o = O.System.quartz(*cells, sigma_ang=sigma, seed=seed)
    lo, hi = o.gershgorin()
    a, b = (hi - lo) / 2, (hi + lo) / 2
    cores = host_cores()
    # grow the sample until it takes at least 0.6 target_s (or covers every column)
    ncol = cores
    while True:
        t0 = time.perf_counter()
        o.diag_moments(a, b, M + 1, 0, ncol, nthreads=cores)
        t = time.perf_counter() - t0
        if t >= 0.6 * target_s or ncol >= o.norb:
            break
        ncol = min(o.norb, max(2 * ncol, int(ncol * target_s / max(t, 1e-3)) // cores * cores))
    n = o.norb
    nnz = o.nnzb * 16
    # the oracle's own FP64 traffic: 3 vector streams + CSR per step, and the
    # reference op sequence's (MatMult + 6 BLAS-1 calls: 14 vector streams + CSR)
    own_bytes_per_col = M * (3 * 8.0 * n + nnz * 12.0 + 4.0 * (n + 1))
    ref_bytes_per_col = M * (14 * 8.0 * n + nnz * 12.0 + 4.0 * (n + 1))
    iter_s = t / ncol * n
    return {"value": round(gpu_bytes_per_iter / iter_s / 1e9, 3),
            "unit": "GB/s (the GPU value's algorithmic bytes per SCC iteration / CPU seconds per iteration)",
            "cores": cores, "kind": "port",
            "sample": f"{ncol} of {n} unit-vector columns, M={M}, FP64 CSR SpMV + BLAS-1 per "
                      f"step (getKPE op sequence), {t:.1f} s",
            "scc_iter_s_extrapolated": round(iter_s, 1),
            "own_traffic_gbs": round(own_bytes_per_col * ncol / t / 1e9, 1),
            "reference_op_sequence_gbs": round(ref_bytes_per_col * ncol / t / 1e9, 1),
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def cpu_model():
    """The host CPU's model name (first processor in /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def c0_scc_both(ctx):
    """The 72-atom 2x2x2 cell (sigma 0.05 A, seed 1, M=200): the whole SCC to
    1e-6 e by the oracle (FP64, one host thread, getKPE op sequence) and by
    the device path, with their charge difference."""
    import numpy as np
    import largetb_amd as L
    import oracle as O
    o = O.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    t0 = time.perf_counter()
    r = o.scc(200, tol=1e-6, maxit=300, nthreads=1)
    cpu_s = time.perf_counter() - t0
    s = L.System.quartz(2, 2, 2, sigma_ang=0.05, seed=1)
    ctx.load(s)
    ctx.scc(200, precision=L.FP64, max_iter=1, tol=1e-30, check=False)  # warm-up
    g = ctx.scc(200, precision=L.FP64, tol=1e-6, max_iter=300, use_graph=1)
    return {"cpu_seconds": round(cpu_s, 3), "cpu_iters": r["iters"], "cores": 1,
            "gpu_seconds": round(g["seconds"], 4), "gpu_iters": g["iters"],
            "max_dq_e": float(np.abs(g["q"] - r["q"]).max())}


def measured_traffic(cells, M, tiling, prec_name):
    """HBM bytes per SpMM launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE / WRITE_SIZE scaled by a plain-copy calibration), when this
    run's workload and tiling are the profiled ones; otherwise None."""
    import re
    path = os.path.join(ROOT, "profiles", "cheb_sweep_fp32_quartz12_M512.txt")
    if not (cells == [12, 12, 12] and prec_name == "fp32" and tiling and os.path.exists(path)):
        return None
    want = (4096, 4, 2, 2, 1, 2)
    got = tuple(tiling[k] for k in ("chunk_cols", "cols_per_lane", "schedule", "blocks_per_cu", "nontemporal",
                                    "order"))
    if got != want:
        return None
    m = re.search(r"= ([0-9.]+) GB per launch", open(path).read())
    return float(m.group(1)) * 1e9 if m else None


def copy_calibration(nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size: the reference point that
    FETCH_SIZE / WRITE_SIZE readings of the SpMM kernel are scaled by."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def time_to_solution(args):
    """Both runs converge to |q_out - q_in|_inf < 1e-6 e with linear mixing 0.2
    from inputs resident on the device; the learned run's clock includes
    descriptors and network inference."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd import learned
    s = L.System.quartz(*args.cells, sigma_ang=args.sigma, seed=args.seed)
    ctx = L.Context(0)
    ctx.load(s)
    m, info = learned.load(ctx, args.weights) if args.weights else learned.load(ctx)
    log(f"weights: {info}")
    # tilings for both precisions, and a short warm-up of both code paths
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    t64, _ = ctx.autotune(L.FP64, 3)  # each precision keeps its own tiling
    t32, _ = ctx.autotune(L.FP32_ACC64, 3)
    ctx.scc(args.M, precision=L.FP32_ACC64, max_iter=1, tol=1e-30, check=False)
    ctx.scc(args.M, precision=L.FP64, max_iter=1, tol=1e-30, check=False)
    runs = {}
    log("plain FP64 from the neutral start")
    t0 = time.perf_counter()
    r = ctx.scc(args.M, precision=L.FP64, tol=1e-6, max_iter=400, use_graph=1, verbose=1)
    runs["plain_fp64"] = dict(seconds=time.perf_counter() - t0, iters=r["iters"], q=r["q"],
                              eband_fp=r["eband_fp"], mu=r["mu_ha"])
    log("learned guess + FP32 bulk, FP64 final iteration(s)")
    t0 = time.perf_counter()
    guess = m.predict(s)
    tpred = time.perf_counter() - t0
    r = ctx.scc(args.M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6,
                max_iter=400, use_graph=1, verbose=1)
    runs["learned_mixed"] = dict(seconds=time.perf_counter() - t0, iters=r["iters"], q=r["q"],
                                 eband_fp=r["eband_fp"], mu=r["mu_ha"], fp64_iters=r["fp64_iters"],
                                 predict_s=tpred,
                                 guess_err=float(np.abs(2 * guess - runs["plain_fp64"]["q"]).max()))
    log("learned guess + two-level SCC")
    t0 = time.perf_counter()
    guess = m.predict(s)
    r = ctx.scc(args.M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000,
                use_graph=1, verbose=1, M_lo=TWO_LEVEL_M_LO)
    runs["learned_mixed_two_level"] = dict(seconds=time.perf_counter() - t0, iters=r["iters"], q=r["q"],
                                           eband_fp=r["eband_fp"], mu=r["mu_ha"], fp64_iters=r["fp64_iters"],
                                           lo_iters=r["lo_iters"], M_lo=TWO_LEVEL_M_LO)
    if not args.no_ablation:
        log("neutral start + FP32 bulk, FP64 final iteration(s)")
        t0 = time.perf_counter()
        r = ctx.scc(args.M, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=400,
                    use_graph=1, verbose=1)
        runs["neutral_mixed"] = dict(seconds=time.perf_counter() - t0, iters=r["iters"], q=r["q"],
                                     eband_fp=r["eband_fp"], mu=r["mu_ha"])
        log("neutral start + two-level SCC")
        t0 = time.perf_counter()
        r = ctx.scc(args.M, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000,
                    use_graph=1, verbose=1, M_lo=TWO_LEVEL_M_LO)
        runs["neutral_two_level"] = dict(seconds=time.perf_counter() - t0, iters=r["iters"], q=r["q"],
                                         eband_fp=r["eband_fp"], mu=r["mu_ha"], fp64_iters=r["fp64_iters"],
                                         lo_iters=r["lo_iters"], M_lo=TWO_LEVEL_M_LO)
    ref = runs["plain_fp64"]
    out = {"metric": METRIC, "workload": f"{s.nat}-atom alpha-quartz {'x'.join(map(str, args.cells))}",
           "M": args.M, "sigma_ang": args.sigma, "seed": args.seed, "tiling_fp32": t32, "tiling_fp64": t64}
    for k, v in runs.items():
        d = {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items() if kk != "q"}
        d["max_dq_vs_plain_e"] = float(np.abs(v["q"] - ref["q"]).max())
        d["deband_vs_plain_ev_per_atom"] = abs(v["eband_fp"] - ref["eband_fp"]) * 27.211386 / s.nat
        d["dmu_vs_plain_ev"] = abs(v["mu"] - ref["mu"]) * 27.211386
        out[k] = d
    out["speedup_learned_mixed_vs_plain_fp64"] = round(ref["seconds"] / runs["learned_mixed"]["seconds"], 3)
    out["speedup_learned_mixed_two_level_vs_plain_fp64"] = round(
        ref["seconds"] / runs["learned_mixed_two_level"]["seconds"], 3)
    print(json.dumps(out))
    m.close()
    ctx.close()


def two_level_check(args):
    """The learned guess + two-level SCC on one structure (both tilings tuned,
    both code paths warmed up first), then one plain FP64 iteration started
    from its answer: that iteration's |dq| is the distance of the answer from
    the plain solve's fixed point (the plain solve itself stops within 1e-6 e
    of it).  For cells whose plain solve is too long to repeat."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd import learned
    EV = 27.211386
    s = L.System.quartz(*args.cells, sigma_ang=args.sigma, seed=args.seed)
    ctx = L.Context(0)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    t64, _ = ctx.autotune(L.FP64, 3)
    t32, _ = ctx.autotune(L.FP32_ACC64, 3)
    for p in (L.FP32_ACC64, L.FP64):
        ctx.scc(args.M, precision=p, max_iter=1, tol=1e-30, check=False)
    ctx.scc(args.M, precision=L.FP32_ACC64, max_iter=2, tol=1e-30, check=False, M_lo=TWO_LEVEL_M_LO)
    m, info = learned.load(ctx, args.weights) if args.weights else learned.load(ctx)
    t0 = time.perf_counter()
    guess = m.predict(s)
    tpred = time.perf_counter() - t0
    r = ctx.scc(args.M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000,
                use_graph=1, verbose=1, M_lo=TWO_LEVEL_M_LO)
    secs = time.perf_counter() - t0
    m.close()
    # one plain FP64 iteration from the answer (per-spin populations q_out / 2)
    chk = ctx.scc(args.M, guess=r["pop"], precision=L.FP64, tol=1e-30, max_iter=1, check=False)
    out = {"workload": f"{s.nat}-atom alpha-quartz {'x'.join(map(str, args.cells))}, M={args.M}, "
                       f"sigma {args.sigma} A, seed {args.seed}, linear mixing 0.2, |dq| < 1e-6 e",
           "run": f"learned guess + two-level SCC (M_lo={TWO_LEVEL_M_LO}, FP32 until |dq| < 1e-2 e, FP64 after)",
           "seconds": round(secs, 3), "predict_s": round(tpred, 3), "status": r["status"], "iters": r["iters"],
           "full_order_iters": r["iters"] - r["lo_iters"], "fp64_iters": r["fp64_iters"],
           "lo_iters": r["lo_iters"], "guess_err_vs_answer_e": float(np.abs(2 * guess - r["q"]).max()),
           "plain_fp64_check": {"dq_e": chk["hist"][0], "seconds": round(chk["seconds"], 3),
                                "deband_ev_per_atom": abs(chk["eband_fp"] - r["eband_fp"]) * EV / s.nat,
                                "dmu_ev": abs(chk["mu_ha"] - r["mu_ha"]) * EV},
           "tiling_fp32": t32, "tiling_fp64": t64, "weights": info,
           "command": "python3 -u bench.py " + " ".join(sys.argv[1:])}
    print(json.dumps(out))
    ctx.close()


def tts_leg(ctx, s, args, t32_iter):
    """Time to solution on the bench structure, every run to |dq| < 1e-6 e from
    inputs resident on the device: plain FP64 from the neutral start; the
    learned guess (descriptors + network inference inside the clock) with the
    two-level SCC; then, if the time budget still allows, the learned guess
    with FP32 bulk iterations and an FP64 finish.  Also the FP64 Chebyshev
    kernel's rate.  The leg checks the remaining time budget before each run
    and skips what would not fit (saying so); the plain run's iteration cap
    follows from the budget too."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd import learned
    EV = 27.211386
    out = {"workload": f"{s.nat}-atom alpha-quartz {'x'.join(map(str, args.cells))}, M={args.M}, "
                       f"sigma {args.sigma} A, seed {args.seed}, linear mixing 0.2, |dq| < 1e-6 e"}
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    t64, _ = ctx.autotune(L.FP64, 3)  # the FP32 tiling was tuned for the value leg
    ctx.set_timing(True)
    w = ctx.scc(args.M, precision=L.FP64, max_iter=1, tol=1e-30, check=False)  # warm-up, eager
    st = ctx.kpe_stats()
    ctx.set_timing(False)
    t64_iter = w["seconds"]
    k1 = st["bytes"] / st["seconds"] / 1e9
    out["k1_fp64"] = {"achieved": round(k1, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(k1 / HBM_PEAK_GBS, 4), "launches": st["launches"],
                      "mean_launch_us": round(st["seconds"] / max(st["launches"], 1) * 1e6, 2), "tiling": t64}
    ctx.scc(args.M, precision=L.FP32_ACC64, max_iter=2, tol=1e-30, check=False, M_lo=TWO_LEVEL_M_LO)  # warm-up
    m, info = learned.load(ctx, args.weights) if args.weights else learned.load(ctx)
    # iteration counts measured on this workload: 73 plain; two-level ~3
    # full-order + ~90 low-order; 40 learned + mixed
    two_est = 1.05 * (2 * t64_iter + t32_iter + 95 * t32_iter * TWO_LEVEL_M_LO / args.M + 2.0)
    plain_est = 1.05 * 76 * t64_iter
    mixed_est = 1.05 * (43 * t32_iter + 2.0)
    left = args.time_budget - (time.perf_counter() - T_START)
    out["estimate_s"] = round(plain_est + two_est + mixed_est, 1)
    if plain_est + two_est > left:
        out["skipped"] = (f"estimated {plain_est + two_est:.0f} s (plain + two-level) exceeds the {left:.0f} s left "
                          f"of the {args.time_budget:.0f} s budget")
        m.close()
        return out
    log("time to solution: plain FP64 from the neutral start")
    cap = int((left - two_est) / (1.02 * t64_iter))
    t0 = time.perf_counter()
    r = ctx.scc(args.M, precision=L.FP64, tol=1e-6, max_iter=min(400, cap), use_graph=1, check=False, verbose=1)
    plain = dict(seconds=time.perf_counter() - t0, iters=r["iters"], status=r["status"])
    qp, ep, mp = r["q"], r["eband_fp"], r["mu_ha"]
    out["plain_fp64"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in plain.items()}
    if r["status"] != 0:
        out["skipped"] = f"plain FP64 stopped at its budget cap of {cap} iterations"
        m.close()
        return out

    def record(name, d, rr):
        out[name] = {k: (round(v, 3) if isinstance(v, float) and k != "guess_err_e" else v) for k, v in d.items()}
        out[name]["max_dq_vs_plain_e"] = float(np.abs(rr["q"] - qp).max())
        out[name]["deband_vs_plain_ev_per_atom"] = abs(rr["eband_fp"] - ep) * EV / s.nat
        out[name]["dmu_vs_plain_ev"] = abs(rr["mu_ha"] - mp) * EV
        out[f"speedup_{name}_vs_plain_fp64"] = round(plain["seconds"] / d["seconds"], 3)

    log(f"time to solution: learned guess + two-level SCC (M_lo={TWO_LEVEL_M_LO}, FP32 until |dq| < 1e-2 e)")
    t0 = time.perf_counter()
    guess = m.predict(s)
    tpred = time.perf_counter() - t0
    r = ctx.scc(args.M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-2, tol=1e-6, max_iter=2000,
                use_graph=1, check=False, verbose=1, M_lo=TWO_LEVEL_M_LO)
    record("learned_mixed_two_level",
           dict(seconds=time.perf_counter() - t0, iters=r["iters"], full_order_iters=r["iters"] - r["lo_iters"],
                fp64_iters=r["fp64_iters"], lo_iters=r["lo_iters"], M_lo=TWO_LEVEL_M_LO, status=r["status"],
                predict_s=round(tpred, 3), guess_err_e=float(np.abs(2 * guess - qp).max())), r)
    left = args.time_budget - (time.perf_counter() - T_START)
    if mixed_est > left:
        out["learned_mixed"] = {"skipped": f"estimated {mixed_est:.0f} s exceeds the {left:.0f} s left"}
    else:
        log("time to solution: learned guess + FP32 bulk + FP64 finish")
        t0 = time.perf_counter()
        guess = m.predict(s)
        tpred = time.perf_counter() - t0
        r = ctx.scc(args.M, guess=guess, precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6,
                    max_iter=max(1, min(400, int(left / t32_iter) - 2)), use_graph=1, check=False, verbose=1)
        record("learned_mixed", dict(seconds=time.perf_counter() - t0, iters=r["iters"], fp64_iters=r["fp64_iters"],
                                     status=r["status"], predict_s=round(tpred, 3)), r)
    m.close()
    out["weights"] = info
    return out


def c1_measure(args):
    """The 8x8x8 cell (4,608 atoms, 18,432 orbitals, sigma 0.05 A), M=512:
    every result of the FP32 forms against FP64 -- the KPE density at the
    neutral start and the converged SCC (each charge, E_band, mu) -- the K1
    rate in both precisions, and dense SP2 (FP32 MFMA bulk + 2 FP64
    iterations vs FP64 throughout) with its MFMA rates."""
    import numpy as np
    import largetb_amd as L
    EV = 27.211386
    s = L.System.quartz(8, 8, 8, sigma_ang=0.05, seed=args.seed)
    ctx = L.Context(0)
    ctx.load(s)
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    out = {"workload": f"{s.nat}-atom alpha-quartz 8x8x8, {s.norb} orbitals, sigma 0.05 A, seed {args.seed}",
           "M": 512}
    tilings, kpe = {}, {}
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32_ACC64)):
        tilings[name], _ = ctx.autotune(prec, 3)
        ctx.kpe_moments(512, prec)  # warm-up
        ctx.set_timing(True)
        am, _ = ctx.kpe_moments(512, prec)
        st = ctx.kpe_stats()
        ctx.set_timing(False)
        d = ctx.density(512)
        gbs = st["bytes"] / st["seconds"] / 1e9
        kpe[name] = dict(d, am=am, gbs=gbs)
        out[f"kpe_{name}"] = {"tiling": tilings[name], "achieved_gbs": round(gbs, 1),
                              "frac": round(gbs / HBM_PEAK_GBS, 4), "launches": st["launches"]}
    a, b = kpe["fp64"], kpe["fp32"]
    out["density_fp32_vs_fp64"] = {
        "max_atom_moment_diff": float(np.abs(a["am"] - b["am"]).max()),
        "max_dq_e": float(2 * np.abs(a["pop"] - b["pop"]).max()),
        "deband_ev_per_atom": abs(a["eband"] - b["eband"]) * EV / s.nat,
        "dmu_ev": abs(a["mu_ha"] - b["mu_ha"]) * EV}
    ctx.set_populations(s.neutral())  # SP2 at the same (neutral) start as the densities above
    sp = {}
    for name, kw in (("fp64", dict(max_iter=0, fp64_tail=100)), ("fp32_fp64tail", dict(max_iter=100, fp64_tail=2))):
        ctx.sp2_dense(max_iter=2, fp64_tail=1)  # warm-up
        t0 = time.perf_counter()
        r = ctx.sp2_dense(**kw)
        sp[name] = dict(r, wall=time.perf_counter() - t0)
    for name, r in sp.items():
        peak = 78.6 if name == "fp64" else 157.3
        tf = r["gemm_flops"] / r["gemm_seconds"] / 1e12
        out[f"sp2_{name}"] = {"iters": r["iters"], "fp64_iters": r["fp64_iters"], "seconds": round(r["wall"], 3),
                              "gemm_ms_per_product": round(1e3 * r["gemm_seconds"] / r["iters"], 2),
                              "gemm_tflops_exec": round(tf, 2), "trace_minus_nocc": r["trace"] - s.nelectrons / 2,
                              "idem": r["idem"], "mfma_peak_tflops": peak}
    out["sp2_fp32_vs_fp64"] = {"max_diag_diff": float(np.abs(sp["fp64"]["diag"] - sp["fp32_fp64tail"]["diag"]).max()),
                               "max_dq_e": float(2 * np.abs(sp["fp64"]["pop"] - sp["fp32_fp64tail"]["pop"]).max())}
    out["sp2_fp64_vs_kpe_fp64_max_dq_e"] = float(2 * np.abs(sp["fp64"]["pop"] - a["pop"]).max())
    runs = {}
    ctx.scc(512, precision=L.FP32_ACC64, max_iter=2, tol=1e-30, check=False, M_lo=TWO_LEVEL_M_LO)  # warm-up
    for name, kw in (("plain_fp64", dict(precision=L.FP64, max_iter=400)),
                     ("mixed", dict(precision=L.FP32_ACC64, switch_tol=1e-6, max_iter=400)),
                     ("fp32_only", dict(precision=L.FP32_ACC64, max_iter=150)),
                     ("two_level", dict(precision=L.FP32_ACC64, switch_tol=1e-2, max_iter=2000, M_lo=TWO_LEVEL_M_LO))):
        t0 = time.perf_counter()
        r = ctx.scc(512, tol=1e-6, use_graph=1, check=False, **kw)
        runs[name] = dict(r, wall=time.perf_counter() - t0)
    ref = runs["plain_fp64"]
    for name, r in runs.items():
        out[f"scc_{name}"] = {"status": r["status"], "iters": r["iters"], "fp64_iters": r["fp64_iters"],
                              "seconds": round(r["wall"], 3),
                              "max_dq_vs_plain_e": float(np.abs(r["q"] - ref["q"]).max()),
                              "deband_vs_plain_ev_per_atom": abs(r["eband_fp"] - ref["eband_fp"]) * EV / s.nat,
                              "dmu_vs_plain_ev": abs(r["mu_ha"] - ref["mu_ha"]) * EV}
        if name == "two_level":
            out[f"scc_{name}"].update(lo_iters=r["lo_iters"], M_lo=TWO_LEVEL_M_LO, start="neutral")
    out["command"] = "python3 -u bench.py --c1"
    print(json.dumps(out))
    ctx.close()


def sp2_measure(args):
    """Dense SP2 on the 8x8x8 cell (18,432 orbitals): FP32 MFMA bulk + FP64
    tail, against the KPE density at M=512; MFMA rate over the X^2 GEMMs."""
    import numpy as np
    import largetb_amd as L
    cells = args.cells if args.cells != [12, 12, 12] else [8, 8, 8]
    s = L.System.quartz(*cells, sigma_ang=0.05, seed=args.seed)
    ctx = L.Context(0)
    ctx.load(s)
    ctx.set_populations(None)
    ctx.sp2_dense(max_iter=3, fp64_tail=1)  # warm-up (allocations, code load)
    t0 = time.perf_counter()
    r = ctx.sp2_dense(max_iter=100, fp64_tail=2)
    wall = time.perf_counter() - t0
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    ctx.kpe_moments(512, L.FP64)
    d = ctx.density(512)
    n = s.norb
    out = {"workload": f"dense SP2, {s.nat}-atom alpha-quartz {'x'.join(map(str, cells))}, {n} orbitals",
           "iterations": r["iters"], "fp64_iterations": r["fp64_iters"], "seconds": round(wall, 3),
           "gemm_seconds": round(r["gemm_seconds"], 4),
           "gemm_ms_per_product": round(1e3 * r["gemm_seconds"] / r["iters"], 2),
           # 2 n^3 per X^2 over the time: the rate a full (unsymmetric) product would need
           "gemm_tflops_alg": round(2.0 * n ** 3 * r["iters"] / r["gemm_seconds"] / 1e12, 2),
           "gemm_tflops_exec": round(r["gemm_flops"] / r["gemm_seconds"] / 1e12, 2),
           "trace_minus_nocc": r["trace"] - s.nelectrons / 2, "idem": r["idem"],
           "max_dpop_vs_kpe_M512": float(np.abs(r["pop"] - d["pop"]).max())}
    print(json.dumps(out))
    ctx.close()


def mlp_measure(args):
    """Training-step time of the learned-guess network (MFMA forward and
    backward) on the descriptors of 256 uneven cells (2x2x2 .. 5x5x4 quartz,
    86,274 atoms) and of the same set repeated 8 times (690,192 atoms).
    Targets are synthetic (per-species constants plus noise): only the time is
    measured.  MFMA flops per atom and Adam step: forward 2 (80 64 + 2 64^2),
    backward with the forward recomputed 2 (80 64 + 2 64^2) + 2 (2 64^2) (layer
    gradients) + 2 (96 64 + 2 64^2) (weight gradients, features padded to 96)."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd.learned import batch_structures
    ctx = L.Context(0)
    feats, Zs, seg, qtot = [], [], [0], []
    for l, m_, n, sg, sd in batch_structures(256, 2026):
        s = L.System.quartz(l, m_, n, sigma_ang=sg, seed=sd)
        Z, _ = s.coords()
        feats.append(ctx.descriptors(s))
        Zs.append(Z)
        seg.append(seg[-1] + s.nat)
        qtot.append(s.nelectrons)
    feat, Z = np.concatenate(feats), np.concatenate(Zs)
    flops_atom = 2 * (80 * 64 + 2 * 64 * 64) * 2 + 2 * (2 * 64 * 64) + 2 * (96 * 64 + 2 * 64 * 64)
    rng = np.random.default_rng(args.seed)
    out = {"workload": "learned-guess MLP training steps (80-64-64-64-1 per species, full batch)",
           "mfma_flops_per_atom_step": flops_atom}
    for rep in (1, 8):
        f = np.tile(feat, (rep, 1))
        z = np.tile(Z, rep)
        sgs = np.concatenate([[0]] + [np.array(seg[1:]) + k * seg[-1] for k in range(rep)]).astype(np.int32)
        t = np.where(z == 14, 2.6, 6.7) + 0.02 * rng.standard_normal(len(z))
        m = L.Mlp(ctx, seed=1)
        m.set_data(f, z, sgs, np.tile(qtot, rep), t)
        m.fit_normalization()
        m.adam(5, lr=1e-3)
        t0 = time.perf_counter()
        m.adam(args.steps * 10, lr=1e-3)
        dt = (time.perf_counter() - t0) / (args.steps * 10)
        out[f"atoms_{len(z)}"] = {"ms_per_adam_step": round(dt * 1e3, 3),
                                  "tflops_mfma": round(flops_atom * len(z) / dt / 1e12, 2)}
        m.close()
    out["command"] = "python3 -u bench.py --mlp"
    print(json.dumps(out))
    ctx.close()


def batch_measure(args):
    """Many independent uneven cells in one batched SCC (2x2x2 .. 5x5x4 quartz,
    sigma 0.02-0.15 A from the seed, M=256), FP32 bulk with FP64 finish, every
    cell to |dq| < 1e-6 e; against the same cells solved one at a time."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd.learned import batch_structures
    cells = batch_structures(args.batch, seed=args.seed)
    t0 = time.perf_counter()
    systems = [L.System.quartz(l, m, n, sigma_ang=sg, seed=sd) for l, m, n, sg, sd in cells]
    nat = sum(s.nat for s in systems)
    log(f"{len(systems)} cells, {nat} atoms, built in {time.perf_counter() - t0:.1f} s")
    ctx = L.Context(0)
    b = L.Batch(ctx, systems)
    M = args.M if args.M != 512 else 256
    t0 = time.perf_counter()
    tiling, tune_gbs = b.autotune(L.FP32_ACC64, 3)
    log(f"batch autotune {time.perf_counter() - t0:.2f} s: {tiling} at {tune_gbs:.0f} GB/s")
    kw = dict(precision=L.FP32_ACC64, switch_tol=1e-6, tol=1e-6, max_iter=args.max_iter, check=False)
    b.scc(M, max_iter=1, tol=1e-30, check=False)                       # warm-up, FP32 path
    b.scc(M, precision=L.FP64, max_iter=1, tol=1e-30, check=False)     # and FP64 path
    ctx.set_timing(True)
    eager = b.scc(M, use_graph=0, **kw)
    st = b.stats()
    ctx.set_timing(False)
    log(f"eager: {eager['batch_iters']} batch iterations, {eager['seconds']:.2f} s, status {eager['status']}")
    if args.no_graph:
        graph = eager
    else:
        graph = b.scc(M, use_graph=1, verbose=1, **kw)
        log(f"graph: {graph['batch_iters']} batch iterations, {graph['seconds']:.2f} s")
    # the same batch with the two-level SCC (FP32 structures: low-order
    # iterations at M_lo between full-order ones, plain FP64 finish)
    b.scc(M, max_iter=2, tol=1e-30, check=False, M_lo=TWO_LEVEL_M_LO)  # warm-up of its code path
    two = b.scc(M, use_graph=1, verbose=1, M_lo=TWO_LEVEL_M_LO, **dict(kw, max_iter=max(kw["max_iter"], 2000)))
    log(f"two-level: {two['batch_iters']} batch iterations ({two['lo_iters']} low-order only), "
        f"{two['seconds']:.2f} s")
    qb = b.split(graph["q"])
    qt = b.split(two["q"])
    b.close()
    # the same cells one at a time (single-structure SCC, same settings, graphs)
    nseq = len(systems) if args.batch_seq <= 0 else min(args.batch_seq, len(systems))
    pick = np.unique(np.linspace(0, len(systems) - 1, nseq).round().astype(int))
    tseq, err, eerr, err2, eerr2 = 0.0, 0.0, 0.0, 0.0, 0.0
    for k in pick:
        ctx.load(systems[k])
        t1 = time.perf_counter()
        r = ctx.scc(M, use_graph=1, **kw)
        tseq += time.perf_counter() - t1
        ref = ctx.scc(M, precision=L.FP64, tol=1e-6, max_iter=args.max_iter, use_graph=1, check=False)
        err = max(err, float(np.abs(qb[k] - ref["q"]).max()))
        eerr = max(eerr, abs(graph["eband_fp"][k] - ref["eband_fp"]) * 27.211386 / systems[k].nat)
        err2 = max(err2, float(np.abs(qt[k] - ref["q"]).max()))
        eerr2 = max(eerr2, abs(two["eband_fp"][k] - ref["eband_fp"]) * 27.211386 / systems[k].nat)
    ctx.close()
    seq_all = tseq / len(pick) * len(systems)
    achieved = st["bytes"] / st["seconds"] / 1e9 if st["seconds"] > 0 else None
    out = {"workload": f"batched SCC, {len(systems)} uneven alpha-quartz cells (2x2x2..5x5x4), "
                       f"{nat} atoms, exact-trace KPE M={M}, FP32 bulk + FP64 finish to 1e-6 e",
           "cells": len(systems), "atoms": nat, "M": M, "seed": args.seed,
           "status": graph["status"],
           "batch_iterations": graph["batch_iters"], "fp64_batch_iterations": graph["fp64_iters"],
           "cell_iterations_min_max": [int(graph["iters"].min()), int(graph["iters"].max())],
           "seconds_graph": None if args.no_graph else round(graph["seconds"], 3),
           "seconds_eager": round(eager["seconds"], 3),
           "cells_per_s": round(len(systems) / graph["seconds"], 2),
           "one_at_a_time_seconds": round(seq_all, 3),
           "one_at_a_time_sample": f"{len(pick)} of {len(systems)} cells, scaled to all",
           "speedup_vs_one_at_a_time": round(seq_all / graph["seconds"], 2),
           "max_dq_vs_fp64_e": err, "max_deband_vs_fp64_ev_per_atom": eerr,
           "two_level": {"M_lo": TWO_LEVEL_M_LO, "status": two["status"], "batch_iterations": two["batch_iters"],
                         "low_order_only_batch_iterations": two["lo_iters"],
                         "fp64_batch_iterations": two["fp64_iters"], "seconds_graph": round(two["seconds"], 3),
                         "cells_per_s": round(len(systems) / two["seconds"], 2),
                         "speedup_vs_batch": round(graph["seconds"] / two["seconds"], 2),
                         "speedup_vs_one_at_a_time": round(seq_all / two["seconds"], 2),
                         "max_dq_vs_fp64_e": err2, "max_deband_vs_fp64_ev_per_atom": eerr2},
           "roofline": {"bound": "hbm", "kernel": "k_batch_step", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                        "achieved": round(achieved, 1) if achieved else None,
                        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                        "launches": st["launches"],
                        "mean_launch_us": round(st["seconds"] / max(st["launches"], 1) * 1e6, 2),
                        "tiling": {k: tiling[k] for k in ("blocks_per_cu", "atoms_per_wave", "nontemporal")}}}
    print(json.dumps(out))


def scc_segment(args):
    """One segment of a long SCC solve, resumable: linear mixing keeps no state
    but the mixed populations, so a run cut into segments (each under one
    process) repeats the uninterrupted iteration sequence.  The state file
    carries the populations, iteration counts and the summed loop seconds."""
    import numpy as np
    import largetb_amd as L
    from largetb_amd import learned
    s = L.System.quartz(*args.cells, sigma_ang=args.sigma, seed=args.seed)
    ctx = L.Context(0)
    ctx.load(s)
    st = dict(np.load(args.state, allow_pickle=False)) if args.state and os.path.exists(args.state) else None
    mixed = args.segment == "mixed"
    prec = L.FP32_ACC64 if mixed else L.FP64
    if st is not None and int(st["fp64_iters"]) > 0:
        prec = L.FP64  # a mixed run that already switched stays in FP64
    ctx.set_populations(None)
    lo, hi = ctx.gershgorin()
    ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
    for p in ([L.FP64, L.FP32_ACC64] if mixed else [L.FP64]):  # each precision keeps its own tiling
        tiling, gbs = ctx.autotune(p, 3)
        log(f"autotune {'fp64' if p == L.FP64 else 'fp32'}: {tiling} {gbs:.0f} GB/s")
    ctx.scc(args.M, precision=prec, max_iter=1, tol=1e-30, check=False)  # warm-up, not counted
    extra = 0.0
    if st is not None:
        guess = st["pop"]
    elif args.start == "learned":
        m, info = learned.load(ctx, args.weights) if args.weights else learned.load(ctx)
        t0 = time.perf_counter()
        guess = m.predict(s)
        extra = time.perf_counter() - t0
        m.close()
    else:
        guess = None
    r = ctx.scc(args.M, guess=guess, precision=prec, switch_tol=1e-6 if mixed else 0.0, tol=1e-6,
                max_iter=args.max_iter, use_graph=1, verbose=1, check=False)
    prev = st or dict(iters=0, fp64_iters=0, seconds=0.0, hist=np.zeros(0))
    out = dict(pop=ctx.mixed_populations(), iters=int(prev["iters"]) + r["iters"],
               fp64_iters=int(prev["fp64_iters"]) + (r["fp64_iters"] if mixed else r["iters"]),
               seconds=float(prev["seconds"]) + r["seconds"] + extra,
               hist=np.concatenate([np.asarray(prev["hist"]), r["hist"]]), status=r["status"],
               q=r["q"], eband_fp=r["eband_fp"], mu=r["mu_ha"], nat=s.nat)
    np.savez(args.state_out, **out)
    summary = {k: (v if not isinstance(v, np.ndarray) else None) for k, v in out.items()
               if k not in ("pop", "q", "hist")}
    summary.update(workload=f"{s.nat}-atom alpha-quartz {'x'.join(map(str, args.cells))}, M={args.M}",
                   run=f"{args.start} start, {args.segment}", segment_iters=r["iters"],
                   segment_seconds=round(r["seconds"], 3), last_dq=float(r["dq"]))
    print(json.dumps(summary))
    ctx.close()


def assembly_measure(args):
    """Neighbour list + H0 + Gamma from coordinates: host cell list
    (ltb_system_create) against the device kernels (ltb_load_coords)."""
    import numpy as np
    import largetb_amd as L
    cells = args.cells if args.cells != [12, 12, 12] else [48, 48, 24]
    s = L.System.quartz(*cells, sigma_ang=args.sigma, seed=args.seed)
    Z, xyz = s.coords()
    a, c = 4.913 * 1.88973, 5.405 * 1.88973
    R = np.array([[a, 0, 0], [-a * np.sin(np.pi / 6), a * np.cos(np.pi / 6), 0], [0, 0, c]])
    cell = np.array([cells[0] * R[0], cells[1] * R[1], cells[2] * R[2]])
    t0 = time.perf_counter()
    h = L.System.from_coords(Z, xyz, cell, periodic_mask=7)
    host_s = time.perf_counter() - t0
    ctx = L.Context(0)
    walls = []  # the first of three runs includes loading the kernels
    for _ in range(3):
        t0 = time.perf_counter()
        inner = ctx.load_coords(Z, xyz, cell)
        walls.append(time.perf_counter() - t0)
    rp, col, val = ctx.hamiltonian()
    hrp, hcol, hval = h.hamiltonian()
    same = bool(np.array_equal(rp, hrp) and np.array_equal(col, hcol))
    out = {"workload": f"H0 + Gamma assembly, {s.nat}-atom alpha-quartz {'x'.join(map(str, cells))}",
           "atoms": s.nat, "blocks": int(ctx.nnzb), "host_cell_list_s": round(host_s, 4),
           "device_s": round(min(walls), 4), "device_inner_s": round(inner, 4),
           "speedup": round(host_s / min(walls), 2), "same_pattern": same,
           "max_abs_diff": float(np.abs(val - hval).max())}
    print(json.dumps(out))
    ctx.close()


def compare_states(paths):
    """Errors of a finished run (second file) against the plain FP64 one (first)."""
    import numpy as np
    a, b = (dict(np.load(p, allow_pickle=False)) for p in paths)
    nat = int(a["nat"])
    print(json.dumps({
        "plain": {"iters": int(a["iters"]), "seconds": round(float(a["seconds"]), 2), "status": int(a["status"])},
        "fast": {"iters": int(b["iters"]), "fp64_iters": int(b["fp64_iters"]),
                 "seconds": round(float(b["seconds"]), 2), "status": int(b["status"])},
        "speedup": round(float(a["seconds"]) / float(b["seconds"]), 3),
        "max_dq_e": float(np.abs(a["q"] - b["q"]).max()),
        "deband_ev_per_atom": abs(float(a["eband_fp"]) - float(b["eband_fp"])) * 27.211386 / nat,
        "dmu_ev": abs(float(a["mu"]) - float(b["mu"])) * 27.211386}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, nargs=3, default=[12, 12, 12])
    ap.add_argument("--sigma", type=float, default=0.10)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--M", type=int, default=512)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tiling", type=str, default="",
                    help="explicit tiling (implies --no-autotune): chunk,V,schedule,blocks_per_cu,nt[,order]")
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--weights", default="", help="--tts / --segment: learned-guess weights (.npz) to use")
    ap.add_argument("--sp2", action="store_true", help="dense SP2 measurement (8x8x8 by default)")
    ap.add_argument("--c1", action="store_true",
                    help="8x8x8 report: FP32 forms vs FP64 (KPE density, SCC, SP2) with K1 and MFMA rates")
    ap.add_argument("--tts", action="store_true",
                    help="SCC time-to-solution: plain FP64 neutral start vs learned guess + mixed")
    ap.add_argument("--two-level", action="store_true",
                    help="learned guess + two-level SCC, checked by one plain FP64 iteration from its answer")
    ap.add_argument("--batch", type=int, default=0,
                    help="batched SCC over this many uneven cells (M=256 unless --M is given)")
    ap.add_argument("--no-graph", action="store_true",
                    help="batch mode: eager iterations only (for kernel tracing)")
    ap.add_argument("--batch-seq", type=int, default=0,
                    help="cells re-solved one at a time for the comparison (0: all; fewer: a sample scaled to all)")
    ap.add_argument("--segment", choices=["plain", "mixed"],
                    help="one resumable segment of an SCC solve (plain FP64 or FP32 bulk + FP64 finish)")
    ap.add_argument("--start", choices=["neutral", "learned"], default="neutral")
    ap.add_argument("--state", default="", help="segment: state file to resume from")
    ap.add_argument("--state-out", default="scc_state.npz")
    ap.add_argument("--max-iter", type=int, default=400)
    ap.add_argument("--compare", nargs=2, metavar=("PLAIN", "FAST"),
                    help="errors and speed-up between two finished segment state files")
    ap.add_argument("--mlp", action="store_true",
                    help="learned-guess network: training-step time of the MFMA kernels")
    ap.add_argument("--assembly", action="store_true",
                    help="host vs device neighbour list + H assembly (48x48x24 cells by default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 1 GiB device-to-device copies (counter calibration)")
    ap.add_argument("--no-tts", action="store_true",
                    help="default run: leave out the time-to-solution leg")
    ap.add_argument("--no-light-cone", action="store_true",
                    help="default run: every recursion step visits every atom (counter runs of the full-width "
                         "step)")
    ap.add_argument("--time-budget", type=float, default=550.0,
                    help="default run: seconds since start within which the time-to-solution leg must end")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.compare:
        compare_states(args.compare)
        return
    if args.assembly:
        assembly_measure(args)
        return
    if args.mlp:
        mlp_measure(args)
        return
    if args.segment:
        scc_segment(args)
        return
    if args.tts:
        time_to_solution(args)
        return
    if args.two_level:
        two_level_check(args)
        return
    if args.sp2:
        sp2_measure(args)
        return
    if args.c1:
        c1_measure(args)
        return
    if args.batch:
        batch_measure(args)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is design-only "
              "(column blocks per GPU + an all-reduce of the moments), so --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)
    import largetb_amd as L
    prec = L.FP32_ACC64 if args.precision == "fp32" else L.FP64
    t0 = time.perf_counter()
    s = L.System.quartz(*args.cells, sigma_ang=args.sigma, seed=args.seed)
    log(f"structure {args.cells}: {s.nat} atoms, {s.nnzb} blocks, {time.perf_counter() - t0:.2f} s")
    ctx = L.Context(0)
    ctx.load(s)
    if args.no_light_cone:
        ctx.set_light_cone(False)
    tiling = None
    if args.tiling:
        vals = [int(x) for x in args.tiling.split(",")]
        ch, v, sch, bpc, nt = vals[:5]
        order = vals[5] if len(vals) > 5 else 0
        ctx.set_tiling(chunk_cols=ch, cols_per_lane=v, schedule=sch, blocks_per_cu=bpc, nontemporal=nt,
                       order=order)
        tiling = ctx.get_tiling(prec)
    elif not args.no_autotune:
        ctx.set_populations(None)
        lo, hi = ctx.gershgorin()
        ctx.rescale((hi - lo) / 2, (hi + lo) / 2)
        t0 = time.perf_counter()
        tiling, gbs = ctx.autotune(prec, trial_steps=3)
        log(f"autotune {time.perf_counter() - t0:.2f} s: {tiling} at {gbs:.0f} GB/s")
    else:
        tiling = ctx.get_tiling(prec)
    if args.warmup > 0:
        r = ctx.scc(args.M, precision=prec, max_iter=args.warmup, tol=1e-30, check=False)
        log(f"warmup {args.warmup} iteration(s): {r['seconds']:.3f} s")
    ctx.set_timing(True)
    # every C-ABI call returns after its stream has drained, so the wall clock
    # around ltb_scc_run brackets exactly the timed device work; HIP events
    # around every step launch of its first iteration give the kernel time
    t0 = time.perf_counter()
    r = ctx.scc(args.M, precision=prec, max_iter=args.steps, tol=1e-30, check=False)
    wall = time.perf_counter() - t0
    st = ctx.kpe_stats()
    log(f"timed {args.steps} iteration(s): {wall:.3f} s, dq={r['dq']:.3e}")
    alg_bytes_iter = st["bytes"]
    traffic = measured_traffic(args.cells, args.M, tiling, args.precision)
    value = alg_bytes_iter * args.steps / wall / 1e9
    achieved = st["bytes"] / st["seconds"] / 1e9
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        # diagonal moments [T_m]_jj, m = 0..M, of every unit vector j, per second of SCC iteration
        "moments_per_s": round((args.M + 1) * s.norb * args.steps / wall, 1),
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32" if prec == L.FP32_ACC64 else "f64",
        "data": "synthetic alpha-quartz, Gaussian displacements from the seed",
        "config": {"workload": f"SCC iteration, {s.nat}-atom alpha-quartz "
                               f"{'x'.join(map(str, args.cells))}, exact-trace KPE M={args.M}",
                   "atoms": s.nat, "orbitals": s.norb, "M": args.M, "sigma_ang": args.sigma,
                   "seed": args.seed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_cheb_sweep" if tiling and tiling.get("schedule") == 2 else "k_cheb_blocked",
                     "launches": st["launches"],
                     "mean_launch_us": round(st["seconds"] / max(st["launches"], 1) * 1e6, 2),
                     "tiling": tiling},
    }
    if traffic is not None:  # the launch the counters measured
        out["roofline"]["traffic_launch"] = "full-width general step (every atom visited), 3.062 GB algorithmic"
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cells, args.sigma, args.seed, args.M, alg_bytes_iter)
        out["cpu_baseline"]["c0_full_scc"] = c0_scc_both(ctx)
    if not args.no_tts and prec == L.FP32_ACC64:
        ctx.load(s)
        out["time_to_solution"] = tts_leg(ctx, s, args, wall / args.steps)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
