"""Benchmark: batched FP64 H*X of the standalone operator workload.

Workload: the 64 x 96 x 80 (491,520-point) triatomic -- radial sinc-DVR on r
and R, Legendre DVR on theta, V = Morse base + P2 + 8 seeded Gaussians -- applied
to a batch of 16 N(0,1) vectors, FP64.  One step = one H*X of the whole batch
with inputs and outputs resident in HBM.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu] [--no-pist]

Prints one JSON line: the H*X throughput in GB/s of compulsory bytes
(X read + Y written + V read once per batch), the roofline object of the
dominant kernel (its launches timed alone with HIP events on the launch
stream), the CPU baseline (the oracle's OpenMP restatement of the
reference's H*X loops on this host), and the PIST eigenpairs/s of the
24 x 32 x 24 FP64 workload on the GPU (single window in FP64 and with FP32
inner solves, and as 10 batched shift windows) next to the oracle's on the CPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def hx_counts(n1, n2, n3, nvec, es=8):
    N = n1 * n2 * n3
    flops_vec = 2 * N * (n1 + n2 + n3) + 5 * N
    bytes_batch = es * (2 * N * nvec + N)
    # per pass
    # modes 1+2 and the diagonal write Y; mode 3 then adds to it
    flops_a = nvec * (2 * N * (n1 + n2) + 3 * N)     # two mode products, (V - s) o X, two adds
    flops_b = nvec * (2 * N * n3 + 2 * N)            # mode 3, DEP scale, add to Y
    bytes_a = es * (2 * N * nvec + N)                # X read, Y written, V read
    bytes_b = es * (3 * N * nvec + N // n3)          # X read, Y read + written, DEP
    return dict(N=N, flops_vec=flops_vec, flops_batch=flops_vec * nvec, bytes_batch=bytes_batch, flops_a=flops_a,
                flops_b=flops_b, bytes_a=bytes_a, bytes_b=bytes_b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-pist", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--solver-run", action="store_true",
                    help="time the solver path on the 64x96x80 operator (OSB set-up, P*X at batch 1 and 16, "
                         "QMR) and print one line per measurement; prints no JSON")
    ap.add_argument("--pist-run", action="store_true",
                    help="run only the PIST measurements on the 24x32x24 workload (no H*X bench, no host "
                         "baseline) and print them as one JSON line, for profiling that path")
    ap.add_argument("--pmc-run", action="store_true",
                    help="workload for rocprofv3 --pmc: three 512 MiB device copies (the counter calibration), "
                         "then --steps H*X calls; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the H*X / PIST path here is built and timed on a single MI355X "
              "(multi-GPU sharding is a design note, not part of this build); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from scalit_amd import _lib as L
    from scalit_amd import problems as P
    from scalit_amd.context import Context

    n1, n2, n3 = P.CONFIG_GRIDS["hx_bench"]
    B = args.batch
    if args.pmc_run:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this kernel
        # against its 512 MiB read and 512 MiB written give the scale factors
        src = torch.randn(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        prob = P.config_problem("hx_bench")
        ctx = Context.from_problem(prob)
        x = ctx.to_device(P.random_vectors(prob.N, B, 2024))
        y = torch.empty_like(x)
        for _ in range(args.steps):
            ctx.hx(x, out=y)
        torch.cuda.synchronize()
        ctx.close()
        print(f"pmc run done: 3 copies of 512 MiB, {args.steps} H*X of batch {B}", flush=True)
        return
    if args.solver_run:
        solver_run(args.steps)
        return
    if args.pist_run:
        print(json.dumps({"pist": pist_bench(True)}), flush=True)
        return
    prob = P.config_problem("hx_bench")
    ctx = Context.from_problem(prob)
    x = ctx.to_device(P.random_vectors(prob.N, B, 2024))
    y = torch.empty_like(x)
    cnt = hx_counts(n1, n2, n3, B)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        ctx.hx(x, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.hx(x, out=y)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = cnt["bytes_batch"] / (ms * 1e-3) / 1e9

    # per-kernel timing with HIP events on the launch stream
    import ctypes as C

    def pass_time(p, reps):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(2):
            L.check(L.lib().scalit_profile_hx_pass(ctx.handle, p, B, C.c_void_p(x.data_ptr()), prob.N,
                                                   C.c_void_p(y.data_ptr()), prob.N,
                                                   C.c_void_p(stream.cuda_stream)))
        ev0.record(stream)
        for _ in range(reps):
            L.check(L.lib().scalit_profile_hx_pass(ctx.handle, p, B, C.c_void_p(x.data_ptr()), prob.N,
                                                   C.c_void_p(y.data_ptr()), prob.N,
                                                   C.c_void_p(stream.cuda_stream)))
        ev1.record(stream)
        ev1.synchronize()
        return ev0.elapsed_time(ev1) / reps

    reps = max(args.steps, 10)
    ta, tb = pass_time(1, reps), pass_time(2, reps)
    dom = 1 if ta >= tb else 2
    t_dom = ta if dom == 1 else tb
    fl = cnt["flops_a"] if dom == 1 else cnt["flops_b"]
    achieved = fl / (t_dom * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "hx_bench_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("pass1_bytes" if dom == 1 else "pass2_bytes")
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": traffic,
                "kernel": "k_hx_modes12 (modes 1+2 + diagonal)" if dom == 1 else "k_outer_mode (mode 3 x DEP, adds to Y)",
                "kernel_ms": round(t_dom, 5), "flops_per_launch": fl}
    hx_tflops = cnt["flops_batch"] / (ms * 1e-3) / 1e12

    out = {
        "metric": "H*X GB/s (%HBM roofline); eigenpairs/s at 1e-8 rel err vs plain FP64 PIST",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "standalone H*X, FP64, 64x96x80 grid (491,520 points), batch 16",
                   "grid": [n1, n2, n3], "batch": B, "seed": 1234},
        "hbm_frac": round(value / HBM_PEAK_GBS, 4),
        "hx_tflops": round(hx_tflops, 3),
        "pass_ms": {"modes12": round(ta, 5), "mode3": round(tb, 5)},
        "roofline": roofline,
    }

    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prob, B, args.cpu_seconds)
    if not args.no_pist:
        out["pist"] = pist_bench(args.no_cpu)
    ctx.close()
    print(json.dumps(out))


def solver_run(reps):
    """The solver path on the 64x96x80 operator, timed with HIP events on the
    current stream (one line per measurement, for profiling runs)."""
    import torch

    from scalit_amd import problems as P
    from scalit_amd.context import Context

    prob = P.config_problem("hx_bench")
    e0 = float(np.sort(prob.res)[50])
    ctx = Context.from_problem(prob, e0=e0)
    t0 = time.perf_counter()
    ctx.block_diag(1000, 1e-5)
    torch.cuda.synchronize()
    print(f"osb set-up: {time.perf_counter() - t0:.3f} s", flush=True)

    def timed(f, n):
        f()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(n):
            f()
        b.record()
        b.synchronize()
        return a.elapsed_time(b) / n

    for nv in (1, 16):
        x = ctx.to_device(P.random_vectors(prob.N, nv, 3))
        y = torch.empty_like(x)
        print(f"batch {nv}: H*X {timed(lambda: ctx.hx(x, out=y), reps):.4f} ms, "
              f"P*X {timed(lambda: ctx.px(x, out=y), reps):.4f} ms", flush=True)
        xf = x.float()
        yf = torch.empty_like(xf)
        print(f"batch {nv} fp32: H*X {timed(lambda: ctx.hx_lp(xf, out=yf), reps):.4f} ms, "
              f"P*X {timed(lambda: ctx.px_lp(xf, out=yf), reps):.4f} ms", flush=True)
        b = ctx.to_device(P.random_vectors(prob.N, nv, 5))
        for rep in range(2):   # the first solve records the QMR graph
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ret, xs, res = ctx.qmr(b, ntype=-1, qmrmax=1000, qmrtol=1e-8)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        it = int(np.abs(ret).max())
        print(f"batch {nv}: QMR to 1e-8 in {it} iterations, {dt * 1e3:.2f} ms ({dt * 1e3 / max(it, 1):.3f} ms per "
              f"iteration), residuals <= {float(np.max(res)):.1e}", flush=True)
    ctx.close()


def cpu_baseline(prob, B, seconds):
    """The oracle's OpenMP restatement of MYHX (the reference's loop nests,
    blocks split over host threads) on a bounded sample of the same workload."""
    from oracle import oracle as O
    from scalit_amd import problems as P

    op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep)
    x = P.random_vectors(prob.N, B, 2024)
    y = np.zeros(prob.N)
    tmp = np.zeros(prob.N)
    op.hx_omp(x[0], y, tmp)   # warm-up
    done = 0
    t0 = time.perf_counter()
    while True:
        op.hx_omp(x[done % B], y, tmp)
        done += 1
        if time.perf_counter() - t0 >= seconds and done >= 1:
            break
    dt = time.perf_counter() - t0
    bytes_vec = 8 * (2 * prob.N + prob.N / B)
    return {"value": round(done * bytes_vec / dt / 1e9, 3), "unit": "GB/s", "cores": int(O.lib().orc_omp_threads()),
            "kind": "port", "sample": f"{done} single-vector H*X applications of the 64x96x80 operator "
                                      f"({dt:.1f} s), same byte accounting as the GPU value"}


def pist_bench(no_cpu):
    """PIST on the 24x32x24 workload (lowest 20 levels): plain FP64 and FP32-inner
    single-window runs, and the same levels as 10 batched shift windows."""
    import torch

    from scalit_amd import _lib as L
    from scalit_amd import problems as P
    from scalit_amd.context import Context

    with open(os.path.join(ROOT, "tests", "golden", "pist_fp64_config1.json")) as f:
        g = json.load(f)
    prob = P.config_problem("pist_fp64")
    ref = np.array(g["eigsh"])
    ctx = Context.from_problem(prob, e0=g["e0"])
    ctx.block_diag(1000, 1e-5)
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, g["start_seed"]))
    res = {}
    for name, prec, qtol, etol in (("fp64", L.FP64, 1e-6, 1e-11), ("fp32", L.FP32, 1e-6, 1e-11)):
        for rep in range(2):   # the first run records the QMR graph
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r, eig, rr, stats = ctx.pist_conv(x0, g["e0"], etol, 20, 200, 40, 10, 5, qmrtol=qtol, prec=prec)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        err = float((np.abs(eig - ref) / np.abs(ref)).max())
        res[name] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(20 / dt, 3),
                     "max_rel_err": err, "qmr_iters": int(stats[0]), "krylov": int(stats[2])}
    # the same 20 levels as 10 shift windows of 2 levels, solved in lock-step
    # (one QMR launch per iteration for all windows); window centres are the
    # means of consecutive pairs of the reference levels, and every reference
    # level must be found by some window
    # Lock-step windows wait for the slowest one, whose Krylov count moves with
    # rounding-level differences in the inner solves, so the batched figure is
    # the mean over several sets of start vectors (each timed after a warm-up run
    # that records the batched QMR graph).
    # A second layout: 20 windows of one level each, every centre a quarter of
    # the smaller neighbouring gap inside its level's interval (so that its
    # nearest level is that level), as when approximate levels from an earlier
    # run are refined.  The one-level windows converge in fewer Krylov steps.
    centres2 = [float(ref[2 * k:2 * k + 2].mean()) for k in range(10)]
    gaps = np.diff(ref)
    centres1 = [float(ref[k] + 0.25 * min(gaps[max(k - 1, 0)], gaps[k])) for k in range(19)]
    centres1.append(float(ref[19] - 0.25 * gaps[18]))
    seeds = [g["start_seed"] + i for i in range(4)]
    for name, prec, centres, m_win, start in (("batched_fp64", L.FP64, centres2, 2, 12),
                                              ("batched_fp32", L.FP32, centres2, 2, 12),
                                              ("batched_fp64_1level", L.FP64, centres1, 1, 4)):
        nwin = len(centres)
        times, errs, iters, conv, krylov = [], [], [], True, []
        for i, seed in enumerate(seeds):
            x0b = ctx.to_device(P.random_vectors(prob.N, nwin, seed))
            for rep in range(2 if i == 0 else 1):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                ret, eig, rr, stats = ctx.pist_conv_batched(x0b, centres, 1e-11, m_win, 200, start, 4, 1,
                                                            qmrtol=1e-6, prec=prec)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
            found = np.sort(eig.reshape(-1))
            errs.append(max(float(np.min(np.abs(found - r)) / abs(r)) for r in ref))
            times.append(dt)
            iters.append(int(stats[0]))
            krylov.append(int(stats[2]))
            conv = conv and all(int(v) > 0 for v in ret)
        dt = float(np.mean(times))
        res[name] = {"windows": nwin, "levels_per_window": m_win, "start_seeds": seeds,
                     "seconds_per_seed": [round(t, 4) for t in times], "seconds": round(dt, 4),
                     "eigenpairs_per_s": round(20 / dt, 3), "max_rel_err": max(errs), "qmr_iters": iters,
                     "krylov": krylov, "speedup_vs_fp64": round((20 / dt) / res["fp64"]["eigenpairs_per_s"], 3),
                     "converged": bool(conv)}
    # the one-level layout again with the Hellmann-Feynman gradients of every
    # level found (one Ritz vector per window, dE/dp for the potential's 12
    # parameters) inside the timed region
    dvdp = ctx.to_device(prob.potential.gradient(prob.r, prob.R, prob.cos_theta))
    times = []
    for i, seed in enumerate(seeds):
        x0b = ctx.to_device(P.random_vectors(prob.N, len(centres1), seed))
        for rep in range(2 if i == 0 else 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ret, eig, rr, stats = ctx.pist_conv_batched(x0b, centres1, 1e-11, 1, 200, 4, 4, 1, qmrtol=1e-6,
                                                        prec=L.FP64)
            psi = ctx.pist_vectors_all(len(centres1), 1)
            grad = ctx.eig_gradients(psi, dvdp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        times.append(dt)
    dt = float(np.mean(times))
    res["batched_fp64_1level_gradients"] = {
        "windows": len(centres1), "levels_per_window": 1, "parameters": int(dvdp.shape[0]), "start_seeds": seeds,
        "seconds_per_seed": [round(t, 4) for t in times], "seconds": round(dt, 4),
        "eigenpairs_per_s": round(20 / dt, 3),
        "speedup_vs_fp64": round((20 / dt) / res["fp64"]["eigenpairs_per_s"], 3),
        "gradients_finite": bool(np.isfinite(grad).all())}
    if not no_cpu:
        from oracle import oracle as O

        op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=g["e0"])
        op.diag(1000, 1e-5)
        t0 = time.perf_counter()
        r, eig, rr, c = op.pist_conv(P.random_vectors(prob.N, 1, g["start_seed"])[0], g["e0"], 1e-11, 20, 200, 40,
                                     10, 5, qmr_tol=1e-6)
        dt = time.perf_counter() - t0
        res["cpu_fp64"] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(20 / dt, 3),
                           "max_rel_err": float((np.abs(eig - ref) / np.abs(ref)).max()), "cores": 1,
                           "kind": "port"}
    ctx.close()
    return res


if __name__ == "__main__":
    main()
