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
import subprocess
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
    ap.add_argument("--config4", choices=["a", "ab", "abc"], default="a",
                    help="grids of the uniform-window layout: 48x64x40 (a), 80x96x64 (b), 128x128x96 (c); "
                         "the default runs the first (about 10 s), abc takes minutes")
    ap.add_argument("--solver-run", action="store_true",
                    help="time the solver path on the 64x96x80 operator (OSB set-up, P*X at batch 1 and 16, "
                         "QMR) and print one line per measurement; prints no JSON")
    ap.add_argument("--pist-run", action="store_true",
                    help="run only the PIST measurements on the 24x32x24 workload (no H*X bench, no host "
                         "baseline) and print them as one JSON line, for profiling that path")
    ap.add_argument("--multi-run", action="store_true",
                    help="run only the multi-problem QMR measurement on the three config-4 grids and print it as "
                         "one JSON line")
    ap.add_argument("--complex-run", action="store_true",
                    help="run only the resonance and CRP measurements on the 24x32x24 operator with an absorbing "
                         "potential and print them as one JSON line, for profiling that path")
    ap.add_argument("--pmc-run", action="store_true",
                    help="workload for rocprofv3 --pmc: three 512 MiB device copies (the counter calibration), "
                         "then --steps H*X calls, the OSB set-up and --steps P*X calls; prints no JSON")
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
        # then the OSB set-up and P*X (TOSB) of the same batch
        ctx.block_diag(1000, 1e-5)
        for _ in range(args.steps):
            ctx.px(x, out=y)
        torch.cuda.synchronize()
        ctx.close()
        print(f"pmc run done: 3 copies of 512 MiB, {args.steps} H*X and {args.steps} P*X of batch {B}", flush=True)
        return
    if args.solver_run:
        solver_run(args.steps)
        return
    if args.pist_run:
        print(json.dumps({"pist": pist_bench(True)}), flush=True)
        return
    if args.multi_run:
        print(json.dumps({"qmr_multi": qmr_multi_bench()}), flush=True)
        return
    if args.complex_run:
        print(json.dumps({"absorbing_potential_24x32x24": complex_bench(),
                          "complex_levels": complex_levels_bench()}), flush=True)
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
        out["gradients_64x96x80"] = gradients_bench(ctx, prob)
        out["uniform_windows"] = config4_bench(["batched_" + c for c in args.config4])
        out["absorbing_potential_24x32x24"] = complex_bench()
        out["complex_levels"] = complex_levels_bench()
        out["qmr_multi"] = qmr_multi_bench()
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
    # the windowed kernel TOSBW: the 64 / 256 OSB states nearest E0, batch 16
    from scalit_amd import _lib as L

    x = ctx.to_device(P.random_vectors(prob.N, 16, 3))
    y = torch.empty_like(x)
    b = ctx.to_device(P.random_vectors(prob.N, 16, 5))
    for cnt in (64, 256):
        t0 = time.perf_counter()
        ctx.set_osbw(e0, 1e-4, cnt)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        tp = timed(lambda: ctx.px(x, nosb=L.TOSBW, out=y), reps)
        ctx.set_osb(L.TOSBW)
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ret, xs, res = ctx.qmr(b, ntype=-1, qmrmax=1000, qmrtol=1e-8)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        ctx.set_osb(L.TOSB)
        it = int(np.abs(ret).max())
        print(f"TOSBW window {cnt}: set-up {ts * 1e3:.1f} ms, batch 16 P*X {tp:.4f} ms, QMR to 1e-8 in {it} "
              f"iterations, {dt * 1e3:.2f} ms, residuals <= {float(np.max(res)):.1e}", flush=True)
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
                                      f"({dt:.1f} s), same byte accounting as the GPU value",
            "host": host_info()}


def host_info():
    """Host cores and CPU model of the machine the baseline runs on."""
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except Exception:   # lscpu missing: leave the model empty
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
    except Exception:
        nproc = os.cpu_count()
    return {"nproc": nproc, "cpu_model": model}


def gradients_bench(ctx, prob):
    """The adjoint-gradient workload: the lowest 50 levels of the 64x96x80
    operator by one FP64 PIST window (ETOL 1e-10, QMR 1e-6), their Ritz
    vectors and dE_k/dp_j for the 12 potential parameters; timed with and
    without the vectors and gradients (OSB set-up outside the timed region)."""
    import torch

    from scalit_amd import problems as P

    ctx.block_diag(1000, 1e-5)
    e0 = float(prob.res.min())
    ctx.set_shift(e0)
    ctx.set_pist_orth(2)   # a Krylov space of ~300 vectors: one MGS pass loses orthogonality
    x0 = ctx.to_device(P.random_vectors(prob.N, 1, 91))
    dvdp = ctx.to_device(prob.potential.gradient(prob.r, prob.R, prob.cos_theta))
    out = {}
    for with_grad in (False, True):
        for rep in range(2):   # the first run records the QMR graph
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r, eig, rr, stats = ctx.pist_conv(x0, e0, 1e-10, 50, 400, 150, 20, 10, qmrtol=1e-6)
            if with_grad:
                grad = ctx.eig_gradients(ctx.pist_vectors(0, 50), dvdp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        key = "pist_and_gradients" if with_grad else "pist_only"
        out[key] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(50 / dt, 3),
                    "qmr_iters": int(stats[0]), "krylov": int(stats[2])}
        if with_grad:
            out[key]["parameters"] = int(grad.shape[1])
            out[key]["gradients_finite"] = bool(np.isfinite(grad).all())
    psi = ctx.pist_vectors(0, 50)
    rr = ctx.hx(psi) - torch.as_tensor(eig, device=psi.device)[:, None] * psi
    out["max_ritz_residual"] = float((rr * rr).sum(dim=1).sqrt().max())
    out["orthogonalisation_passes"] = 2
    ctx.set_pist_orth(1)
    return out


def qmr_multi_bench():
    """FP64 QMR on several operators of uneven grids with 4 right-hand sides
    each, shifts between the 4th and 5th OSB levels of each, to 1e-8: all
    problems in lock step with one graph launch per iteration
    (scalit_qmr_multi) against the same solves one problem after another (the
    one-kernel-per-step path); on the three config-4 grids (48x64x40,
    80x96x64, 128x128x96) and on three small grids (24x32x24, 20x28x16,
    16x20x12), where single solves are launch-bound.  OSB set-ups outside the
    timed regions; the second of two runs is timed."""
    import torch

    from scalit_amd import problems as P
    from scalit_amd.context import Context

    def one(probs):
        runs = []
        for k, prob in enumerate(probs):
            ctx = Context.from_problem(prob)
            ctx.block_diag(1000, 1e-5)
            ctx.set_qmr_persistent(0)
            _, eig0 = ctx.get_precond()
            srt = np.sort(eig0)
            e0 = float(0.5 * (srt[3] + srt[4]))
            shifts = [e0 + 2e-5 * j for j in range(4)]
            runs.append((prob, ctx, ctx.to_device(P.random_vectors(prob.N, 4, 60 + k)), shifts))
        problems = [(ctx, b, sh) for _, ctx, b, sh in runs]
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            multi = Context.qmr_multi(problems, ntype=-1, qmrmax=300, qmrtol=1e-8)
            t_multi = time.perf_counter() - t0
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            alone = [ctx.qmr(b, ntype=-1, qmrmax=300, qmrtol=1e-8, shifts=sh) for _, ctx, b, sh in runs]
            t_alone = time.perf_counter() - t0
        same = all(bool((m[1] == a[1]).all()) and np.array_equal(m[0], np.array(a[0])) for m, a in zip(multi, alone))
        out = {"grids": [list(p.sN) for p, _, _, _ in runs], "rhs_per_problem": 4,
               "qmr_iterations": [int(np.abs(m[0]).max()) for m in multi],
               "converged": [bool((m[0] > 0).all()) for m in multi],
               "seconds_one_graph_per_iteration": round(t_multi, 4),
               "seconds_one_problem_after_another": round(t_alone, 4),
               "speedup": round(t_alone / t_multi, 3), "bit_identical": same}
        for _, ctx, _, _ in runs:
            ctx.close()
        return out

    big = one([P.config_problem(n) for n in ("batched_a", "batched_b", "batched_c")])
    small = one([P.build_problem(*g, P.jittered_potential(21 + i), seed=21 + i)
                 for i, g in enumerate([(24, 32, 24), (20, 28, 16), (16, 20, 12)])])
    return {"config4_grids": big, "small_grids": small}


def complex_bench():
    """The resonance and CRP paths on the 24x32x24 operator with an absorbing
    potential in R (strength 0.01 hartree beyond 70% of the R range): the 4
    resonance energies whose real parts are nearest a level (PIST_CONV_SX,
    ETOL 1e-12, QMR 1e-10) and N(E) at 3 energies from 10 Lanczos steps
    (product absorbing potential in r), OSB set-up outside the timed region."""
    import torch

    from scalit_amd import _lib as L
    from scalit_amd import problems as P
    from scalit_amd.context import Context

    prob = P.config_problem("pist_fp64")
    ctx = Context.from_problem(prob, e0=0.0)
    ctx.block_diag(1000, 1e-5)
    _, eig0 = ctx.get_precond()
    srt = np.sort(eig0)
    e0 = float(0.5 * (srt[30] + srt[31]))
    ctx.set_shift(e0)
    ap = P.absorbing_potential(prob, 0.01)
    ctx.set_ap(L.AP, ap)
    x0 = ctx.to_device_cx(P.random_vectors(prob.N, 1, 5)[0] + 0j)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r, eig, res, st = ctx.pist_conv_sx(x0, e0, 1e-12, 4, 160, 20, 5, 2, qmr_type=-1, qmrmax=3000, qmrtol=1e-10)
    dt = time.perf_counter() - t0
    out = {"resonance": {"ret": int(r), "seconds": round(dt, 3), "energies": [[float(z.real), float(z.imag)]
                                                                            for z in eig],
                         "qmr_iters": int(st[0]), "solves": int(st[1]), "krylov": int(st[2]),
                         "us_per_qmr_iteration": round(dt * 1e6 / max(1, int(st[0])), 1)}}
    ctx.set_ap(L.APP, P.absorbing_potential(prob, 0.01, 0.7, coord="r"))
    ctx.set_ap(L.APR, ap)
    energies = [e0, e0 + 2e-4, e0 + 4e-4]
    xs = ctx.to_device_cx(np.stack([P.random_vectors(prob.N, 1, 7 + k)[0] + 0j for k in range(3)]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    crp, cst = ctx.crp(energies, 10, xs, qmr_type=-1, qmrmax=3000, qmrtol=1e-10)
    dt = time.perf_counter() - t0
    out["crp"] = {"energies": energies, "crp": [float(v) for v in crp], "seconds": round(dt, 3),
                  "qmr_iters": int(cst[0]), "solves": int(cst[1]),
                  "us_per_qmr_iteration": round(dt * 1e6 / max(1, int(cst[0])), 1)}
    ctx.close()
    return out


def _scx_context(prob, theta=0.05):
    from scalit_amd import problems as P
    from scalit_amd.context import Context

    h0, dep = P.complex_levels(prob, theta, 1e-3, 11)
    ctx = Context(prob.sN, [d is not None for d in prob.dep])
    for l in range(3):
        ctx.set_h0_cx(l + 1, h0[l])
        if dep[l] is not None:
            ctx.set_dep_cx(l + 1, dep[l])
    ctx.set_res(prob.res)
    ctx.set_scx(True)
    return ctx


def complex_levels_bench():
    """Complex level matrices (sCX = T, complex-scaled kinetic matrices,
    theta = 0.05): MYHX_CX of one complex vector on the 64x96x80 grid (HIP
    events around 20 applications after 3 warm-up ones; bytes counted as the
    kernels move them: 40 B per point for the diagonal pass, 48 B per point
    for each level's read of X and read-modify-write of Y), and on 24x32x24
    the sCX OSB set-up and the 4 resonance energies nearest a level
    (PIST_CONV_SX, ETOL 1e-12, QMR 1e-10, absorbing potential in R)."""
    import torch

    from scalit_amd import _lib as L
    from scalit_amd import problems as P

    out = {}
    prob = P.config_problem("hx_bench")
    ctx = _scx_context(prob)
    x = ctx.to_device_cx(P.random_vectors(prob.N, 1, 3)[0] + 1j * P.random_vectors(prob.N, 1, 4)[0])
    for _ in range(3):
        ctx.hx_cx(x, L.TA0)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(20):
        ctx.hx_cx(x, L.TA0)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / 20
    n = prob.N
    nbytes = n * (40 + 3 * 48)
    flops = n * (8 * sum(prob.sN) + 8 + 6)
    out["hx_cx_64x96x80"] = {"ms": round(ms, 4), "GB_per_s": round(nbytes / ms / 1e6, 1),
                             "TFLOP_per_s": round(flops / ms / 1e9, 2), "bytes": nbytes, "flops": flops}
    ctx.close()
    prob = P.config_problem("pist_fp64")
    ctx = _scx_context(prob)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.block_diag(1000, 1e-5)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    _, eig0 = ctx.get_precond()
    srt = np.sort(eig0)
    e0 = float(0.5 * (srt[30] + srt[31]))
    ctx.set_shift(e0)
    ctx.set_ap(L.AP, P.absorbing_potential(prob, 0.01))
    x0 = ctx.to_device_cx(P.random_vectors(prob.N, 1, 5)[0] + 0j)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r, eig, res, st = ctx.pist_conv_sx(x0, e0, 1e-12, 4, 160, 20, 5, 2, qmr_type=-1, qmrmax=3000, qmrtol=1e-10)
    dt = time.perf_counter() - t0
    out["resonance_24x32x24"] = {"setup_seconds": round(setup, 3), "ret": int(r), "seconds": round(dt, 3),
                                 "energies": [[float(z.real), float(z.imag)] for z in eig],
                                 "qmr_iters": int(st[0]), "solves": int(st[1]),
                                 "us_per_qmr_iteration": round(dt * 1e6 / max(1, int(st[0])), 1)}
    ctx.close()
    return out


def config4_bench(names):
    """Batched PIST windows on the configured grids: 32 centres spread
    uniformly over [0, 0.05] hartree, 4 levels each, FP64 (ETOL 1e-10, QMR
    1e-6, at most 3000 QMR iterations per solve), OSB set-up outside the
    timed region.  Windows whose QMR solves do not converge return the
    reference's failure codes and are counted, not hidden."""
    import torch

    from scalit_amd import problems as P
    from scalit_amd.context import Context

    out = {}
    cent = [0.05 * (k + 0.5) / 32 for k in range(32)]
    for name in names:
        prob = P.config_problem(name)
        ctx = Context.from_problem(prob, e0=0.025)
        ctx.block_diag(1000, 1e-5)
        x0 = ctx.to_device(P.random_vectors(prob.N, 32, 5))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ret, eig, rr, st = ctx.pist_conv_batched(x0, cent, 1e-10, 4, 120, 12, 6, 2, qmrtol=1e-6, qmrmax=3000)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ok = ret > 0
        lv = np.sort(eig[ok].reshape(-1))
        dup = int(np.sum(np.diff(lv) <= 1e-9 * np.maximum(1.0, np.abs(lv[1:])))) if lv.size else 0
        out[name] = {"grid": list(prob.sN), "points": int(prob.N), "windows": 32, "levels_per_window": 4,
                     "windows_converged": int(ok.sum()), "highest_converged_centre": float(max(
                         [cent[k] for k in range(32) if ok[k]], default=0.0)),
                     "levels_found": int(lv.size), "duplicated_levels": dup, "seconds": round(dt, 3),
                     "eigenpairs_per_s": round(lv.size / dt, 3), "qmr_iters": int(st[0]), "krylov_max": int(st[2])}
        ctx.close()
    return out


def window_centres(levels):
    """One window per level, centred a quarter of the smaller neighbouring gap
    inside the level's interval, so that its nearest level is that level."""
    lv = np.sort(np.asarray(levels))
    gaps = np.diff(lv)
    c = [float(lv[k] + 0.25 * min(gaps[max(k - 1, 0)], gaps[k])) for k in range(len(lv) - 1)]
    c.append(float(lv[-1] - 0.25 * gaps[-1]))
    return c


def pist_bench(no_cpu):
    """PIST on the 24x32x24 workload (lowest 20 levels, ETOL 1e-11): plain FP64
    and FP32-inner single windows; and one step of a potential-parameter scan
    solved as batched one-level windows placed from the previous step's
    levels, with every level's Ritz vector and Hellmann-Feynman gradients."""
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
        if name == "fp64":
            eig_fp64 = eig.copy()
        err = float((np.abs(eig - ref) / np.abs(ref)).max())
        res[name] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(20 / dt, 3),
                     "max_rel_err": err, "qmr_iters": int(stats[0]), "krylov": int(stats[2]),
                     "us_per_qmr_iteration": round(1e6 * dt / max(int(stats[0]), 1), 1)}
    res["fp64"]["path"] = "persistent (the Krylov steps of each block in one launch)"
    # the same solve with one kernel per QMR step (graph replay) and the
    # reference's MGS_ORTH one projection at a time
    ctx.set_qmr_persistent(0)
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r, eig, rr, stats = ctx.pist_conv(x0, g["e0"], 1e-11, 20, 200, 40, 10, 5, qmrtol=1e-6)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
    ctx.set_qmr_persistent(-1)
    res["fp64_per_step"] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(20 / dt, 3),
                            "max_rel_err": float((np.abs(eig - ref) / np.abs(ref)).max()),
                            "qmr_iters": int(stats[0]), "krylov": int(stats[2]),
                            "us_per_qmr_iteration": round(1e6 * dt / max(int(stats[0]), 1), 1),
                            "path": "one kernel per QMR step, HIP graph replay"}
    res["persistent_speedup"] = round(res["fp64_per_step"]["seconds"] / res["fp64"]["seconds"], 3)
    # Continuation along a potential-parameter scan: the levels of the plain
    # FP64 solve above (this potential) place one window per level for the
    # next scan point, the potential with its 6 Morse parameters moved by a
    # seeded 0.2%.  There the batched one-level windows (20 windows in
    # lock-step), every level's Ritz vector and the Hellmann-Feynman gradients
    # for the 12 parameters are timed against a plain FP64 PIST of the same
    # new operator, which is also the reference for the errors.
    prev = np.sort(eig_fp64)
    pot = prob.potential
    u = P.rng(4321).uniform(-1.0, 1.0, 12)
    p_new = pot.params() * (1.0 + 0.002 * u)
    prob2 = prob.with_potential(pot.with_params(p_new))
    ctx.set_res(prob2.res)
    for rep in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r2, ref2, _, st_plain = ctx.pist_conv(x0, g["e0"], 1e-11, 20, 200, 40, 10, 5, qmrtol=1e-6)
        torch.cuda.synchronize()
        t_plain = time.perf_counter() - t0
    ref2 = np.sort(ref2)
    centres = window_centres(prev)
    dvdp = ctx.to_device(prob2.potential.gradient(prob.r, prob.R, prob.cos_theta))
    seeds = [g["start_seed"] + i for i in range(4)]
    times, errs, iters, krylov = [], [], [], []
    converged, grad_equal = True, True
    for i, seed in enumerate(seeds):
        xs = ctx.to_device(P.random_vectors(prob.N, 20, seed))
        for rep in range(2 if i == 0 else 1):   # the first run records the graphs
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ret, eig, rr, st2 = ctx.pist_conv_batched(xs, centres, 1e-11, 1, 200, 4, 4, 1, qmrtol=1e-6)
            psi = ctx.pist_vectors_all(20, 1)
            grad = ctx.eig_gradients(psi, dvdp)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
        converged = converged and r2 > 0 and bool((ret > 0).all())
        # one window per level: the sorted levels found against the plain
        # solve's, so a level found twice (and another missed) shows as an error
        errs.append(float((np.abs(np.sort(eig.reshape(-1)) - ref2) / np.abs(ref2)).max()))
        # the gradients of the batched call against one window at a time
        one = np.concatenate([ctx.eig_gradients(ctx.pist_vectors(k, 1), dvdp) for k in range(20)])
        grad_equal = grad_equal and bool(np.array_equal(one, grad))
        times.append(dt)
        iters.append(int(st2[0]))
        krylov.append(int(st2[2]))
    ctx.set_res(prob.res)
    dt = float(np.mean(times))
    res["scan_step_batched_gradients"] = {
        "windows": 20, "levels_per_window": 1, "parameters": int(dvdp.shape[0]),
        "placement": "one window per level of the previous scan point (Morse parameters moved by 0.2%)",
        "plain_fp64_seconds": round(t_plain, 4), "plain_fp64_ret": int(r2),
        "start_seeds": seeds, "seconds_per_seed": [round(t, 4) for t in times], "seconds": round(dt, 4),
        "eigenpairs_per_s": round(20 / dt, 3), "max_rel_err_vs_plain_fp64": max(errs), "qmr_iters": iters,
        "krylov": krylov, "speedup_vs_plain_fp64": round(t_plain / dt, 3), "converged": converged,
        "gradients_equal_one_window_at_a_time": grad_equal}
    if not no_cpu:
        from oracle import oracle as O

        op = O.Operator(prob.sN, prob.h0, prob.res, prob.dep, e0=g["e0"])
        op.diag(1000, 1e-5)
        t0 = time.perf_counter()
        r, eig, rr, c = op.pist_conv(P.random_vectors(prob.N, 1, g["start_seed"])[0], g["e0"], 1e-11, 20, 200, 40,
                                     10, 5, qmr_tol=1e-6, threaded=True)
        dt = time.perf_counter() - t0
        res["cpu_fp64"] = {"ret": int(r), "seconds": round(dt, 3), "eigenpairs_per_s": round(20 / dt, 3),
                           "max_rel_err": float((np.abs(eig - ref) / np.abs(ref)).max()),
                           "cores": int(O.lib().orc_omp_threads()), "kind": "port",
                           "note": "the oracle's PIST_CONV with H*X and P*X split over the host threads"}
    ctx.close()
    return res


if __name__ == "__main__":
    main()
