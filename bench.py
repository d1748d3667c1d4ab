"""Benchmark of the hot path on one MI355X.

Step = one variable-density pressure-Poisson solve to ||r||/||b|| <= 1e-10 on
the 2048 x 2048 workload (64 seeded circles, density 1 vs 1000 smoothed over
one cell, zero-mean Gaussian right-hand side, walls): Jacobi-preconditioned CG
in FP64, the reference's arithmetic.  The same run times the mixed-precision
solve of the same system (FP32 Krylov recurrence, FP64 reliable updates of
the true residual) and the multigrid-preconditioned solves, the ELVIRA +
Weymouth-Yue VOF step at 2048^2 and 4096^2 (128 seeded ellipses,
divergence-free Fourier velocity, CFL 0.25), the SP predictor, the whole
two-phase time step at 4096 x 2048 (500 BDF2 steps), the 16-step
checkpointed kinetic-energy adjoint at 4096 x 2048, the projection adjoint,
the plain-method case and the monolithic step, and, on the host cores, the
oracle's FP64 Jacobi-PCG solve of the same system and its ELVIRA + WY step as
the CPU baselines.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
    python bench.py --kernel-sample   # bounded run of the same kernels for
                                      # rocprofv3 counter passes (no JSON)
    python bench.py --summarize DB    # markdown summary of a rocprofv3
                                      # results database (kernel durations,
                                      # and HBM traffic when it holds counters)

The kernel-sample run executes 256 iterations of the mixed solve, 64 of the
FP64 solve and 2 VOF steps on the same workloads, 4 momentum-operator
products and one BiCGSTAB(2) sweep of the 2048^2 predictor, the FP64 and
mixed multigrid-preconditioned solves, then a plain copy of 256 MiB (torch
copy_ of float64, reads and writes 256 MiB) whose FETCH_SIZE / WRITE_SIZE
readings calibrate the counters into bytes.  With --n 4096 the same sample
runs at 4096^2, where the FP32 Jacobi working set (8 arrays, 537 MB) no
longer fits the 256 MB Infinity Cache.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "covof_2d-fortran_amd"))
sys.path.insert(0, ROOT)

METRIC = "Poisson solve s to 1e-10 + VOF cell-updates/s at 2048^2; HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0
# HBM bytes per launch (FETCH_SIZE x 2.0 + WRITE_SIZE, calibrated on a 256 MiB
# copy) measured with rocprofv3 --pmc on `bench.py --kernel-sample` and
# summarised by `bench.py --summarize`; the summary with its commands is
# profiles/poisson_wy_momentum_2048_hbm_traffic.md.
TRAFFIC_2048 = {"k_pcg_a<double,1>": 1.790e8 + 1.007e8, "k_pcg_b<double,1>": 1.008e8 + 3.365e7,
                "k_pcg_a<float,1>": 8.947e7 + 5.036e7, "k_pcg_b<float,1>": 5.047e7 + 1.688e7}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_sample(cv, wl, ctx, b, n, dev):
    import torch
    for precision, iters in ((cv.PRECISION_MIXED, 256), (cv.PRECISION_FP64, 64)):
        x = ctx.new_work()
        r = ctx.poisson_solve(b, x, precond=cv.PRECOND_JACOBI, tol=1e-10, max_iter=iters, precision=precision)
        log(f"sample solve precision={precision}: {r.iters} iterations")
    Cv, uv, vv, dtv = wl.vof_case(n, seed=0)
    vctx = cv.Context(n, n, bc=(cv.BC_PERIODIC,) * 4)
    Cg, ug, vg = (torch.tensor(a, device=dev) for a in (Cv, uv, vv))
    vctx.advect_vof(Cg, ug, vg, dtv, reconstruction_method="ELVIRA", nsteps=2)
    # the momentum operator at 2048^2: 4 plain and 4 Jacobi-divided products
    fm, dtm, bcm, lxm, lym = wl.momentum_case(n, n, seed=0)
    mctx = cv.Context(n, n, lxm, lym, bc=bcm)
    mcfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98)
    F = {k: torch.tensor(fm[k], device=dev) for k in cv.Context.MOMENTUM_FIELDS}
    mctx.momentum_set_fields(mcfg, F, dtm)
    xm = mctx.assemble_rhs()
    for _ in range(4):
        mctx.apply_system_matrix(xm)
    mctx.solve_linear_system(xm, mctx.new_momentum_vector(), cv.LINEAR_BICGSTAB2, cv.PRECOND_JACOBI, 1e-30, 2)
    torch.cuda.synchronize()
    log("momentum sample done")
    # the multigrid-preconditioned solves (FP64, and FP64 CG with the FP32 V-cycle)
    for precision in (cv.PRECISION_FP64, cv.PRECISION_MIXED):
        r = ctx.poisson_solve(b, ctx.new_work(), precond=cv.PRECOND_MG, tol=1e-10, max_iter=100, precision=precision)
        log(f"sample MG solve precision={precision}: {r.iters} iterations")
    a = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
    c = torch.empty_like(a)
    for _ in range(3):
        c.copy_(a)
    torch.cuda.synchronize()
    log("kernel sample done")


def _short(name):
    """Kernel name without its argument list and namespaces."""
    base = name.replace("(anonymous namespace)::", "").split("(")[0]
    return base.replace("void ", "").replace("covof::", "").strip()


def summarize(path, min_us=5.0):
    """Markdown summary of a rocprofv3 results database: per kernel the
    launches, total and average duration over all launches and over the
    launches of at least min_us (the PCG kernels of a captured cycle return
    at once after the stop flag); with counter data (FETCH_SIZE, WRITE_SIZE)
    the HBM bytes per launch, calibrated on the 256 MiB copy of the kernel
    sample (a dispatch whose name holds "copyBuffer")."""
    import sqlite3
    db = sqlite3.connect(path)
    rows = db.execute("select name, end - start, grid_x from kernels").fetchall()
    # a kernel launched on grids of several sizes is listed once per grid
    grids = {}
    for name, _, gx in rows:
        grids.setdefault(_short(name), set()).add(gx)

    def label(name, gx):
        k = _short(name)
        return f"{k} [grid {gx}]" if len(grids.get(k, ())) > 1 else k

    stats = {}
    for name, dur, gx in rows:
        st = stats.setdefault(label(name, gx), [0, 0.0, 0, 0.0])
        st[0] += 1
        st[1] += dur
        if dur >= min_us * 1e3:
            st[2] += 1
            st[3] += dur
    total = sum(v[1] for v in stats.values())
    lines = [f"rocprofv3 database: {os.path.basename(path)}", "",
             f"| kernel | launches | total ms | % | avg us (all) | launches >= {min_us:g} us | avg us (those) |",
             "|---|---|---|---|---|---|---|"]
    for k, (c, tot, cf, tf) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        if tot < 1e-3 * total:
            continue
        lines.append(f"| {k} | {c} | {tot / 1e6:.3f} | {100 * tot / total:.1f} | {tot / c / 1e3:.2f} | {cf} | "
                     + (f"{tf / cf / 1e3:.2f} |" if cf else "- |"))
    try:
        cols = [r[1] for r in db.execute("pragma table_info(counters_collection)")]
        gcol = "grid_x" if "grid_x" in cols else ("grid_size" if "grid_size" in cols else "0")
        pmc = db.execute(f"select kernel_name, counter_name, value, end - start, {gcol} "
                         "from counters_collection").fetchall()
    except sqlite3.OperationalError:
        pmc = []
    if pmc:
        # per kernel and counter: the median over the launches that did full
        # work (at least half the kernel's largest reading; counter collection
        # serialises dispatches, so durations cannot tell early returns apart)
        pgrids = {}
        for name, _, _, _, gx in pmc:
            pgrids.setdefault(_short(name), set()).add(gx)
        vals = {}
        for name, ctr, val, dur, gx in pmc:
            k = _short(name)
            k = f"{k} [grid {gx}]" if len(pgrids[k]) > 1 else k
            vals.setdefault((k, ctr), []).append(val)
        med, nfull = {}, {}
        for k, v in vals.items():
            full = sorted(x for x in v if x >= 0.5 * max(v))
            med[k] = full[len(full) // 2] if full else 0.0
            nfull[k] = len(full)
        cal = {}
        for (k, ctr), v in vals.items():
            # the 256 MiB copies are the largest copy dispatches of the sample
            if "copyBuffer" in k and ctr in ("FETCH_SIZE", "WRITE_SIZE") and max(v) > 0:
                f = 256.0 * 1024 * 1024 / (max(v) * 1024.0)
                cal[ctr] = min(cal.get(ctr, f), f)
        lines += ["", "Counters: median per full-work launch (FETCH_SIZE / WRITE_SIZE in KiB); bytes = KiB x 1024 "
                  "x the 256 MiB copy's calibration " + ", ".join(f"{c} x{f:.3f}" for c, f in sorted(cal.items())),
                  "", "| kernel | counter | launches | value/launch | bytes/launch |", "|---|---|---|---|---|"]
        for (k, ctr), v in sorted(med.items(), key=lambda kv: (kv[0][1], -kv[1])):
            f = cal.get(ctr)
            lines.append(f"| {k} | {ctr} | {nfull[(k, ctr)]} | {v:.6g} | " +
                         (f"{v * 1024.0 * f:.4g} |" if f else "- |"))
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--vof-steps", type=int, default=100)
    ap.add_argument("--sp-steps", type=int, default=500)
    ap.add_argument("--cpu-vof-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--kernel-sample", action="store_true")
    ap.add_argument("--summarize", metavar="DB", help="summarise a rocprofv3 results database and exit")
    ap.add_argument("--min-us", type=float, default=5.0,
                    help="--summarize: launches shorter than this count as early returns")
    args = ap.parse_args()
    if args.summarize:
        print(summarize(args.summarize, args.min_us))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the hot path runs on a single MI355X here "
              "(multi-GPU sharding is described, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import covof_amd as cv
    from covof_amd import _lib as L
    from covof_amd import workloads as wl

    n = args.n
    dev = "cuda"
    # ---------------------------------------------------------------- Poisson
    t0 = time.time()
    case = wl.poisson_case(n, seed=0)
    log(f"poisson case {n}^2 built in {time.time() - t0:.1f}s")
    ctx = cv.Context(n, n, bc=case["bc"])
    T = {k: torch.tensor(case[k], device=dev) for k in ("u", "v", "rho_u", "rho_v")}
    b = ctx.poisson_assemble(T["u"], T["v"], T["rho_u"], T["rho_v"], case["dt"])
    tol = 1e-10
    if args.kernel_sample:
        kernel_sample(cv, wl, ctx, b, n, dev)
        return

    def solve(precision):
        x = ctx.new_work()
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = ctx.poisson_solve(b, x, precond=cv.PRECOND_JACOBI, tol=tol, max_iter=400000, precision=precision)
        torch.cuda.synchronize()
        return x, r, time.perf_counter() - t

    cells = n * n

    def timed(precision, steps, tag_a, tag_b, word):
        """steps timed solves (every 16th iteration's two kernels bracketed by
        HIP events on the solve's stream), the dominant kernel's figures."""
        ctx.profile(16)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        res = []
        for _ in range(steps):
            x = ctx.new_work()
            res.append(ctx.poisson_solve(b, x, precond=cv.PRECOND_JACOBI, tol=tol, max_iter=400000,
                                         precision=precision))
        torch.cuda.synchronize()
        t_solve = (time.perf_counter() - t_start) / steps
        (ms_a, n_a), (ms_b, n_b) = ctx.profile_read(tag_a), ctx.profile_read(tag_b)
        ctx.profile(0)
        assert all(r.converged and r.relres <= tol for r in res), res
        # algorithmic bytes per launch: A reads r, d_prev, x, kx, ky and writes
        # d, Ad, x; B reads r, Ad, a_P and writes r
        bytes_a, bytes_b = 8 * word * cells, 4 * word * cells
        avg_a, avg_b = ms_a / max(n_a, 1), ms_b / max(n_b, 1)
        name = "double" if word == 8 else "float"
        if ms_b >= ms_a:
            dom = (f"k_pcg_b<{name},1>", bytes_b, avg_b, n_b)
        else:
            dom = (f"k_pcg_a<{name},1>", bytes_a, avg_a, n_a)
        iter_gbs = (bytes_a + bytes_b) / ((avg_a + avg_b) * 1e-3) / 1e9
        return x, res[-1], t_solve, dom, iter_gbs

    # the headline: the FP64 Jacobi-PCG solve (the reference's arithmetic)
    for _ in range(args.warmup):
        solve(cv.PRECISION_FP64)
    x64, r64, t64, dom64, gbs64 = timed(cv.PRECISION_FP64, args.steps, L.KERNEL_PCG_A_FP64, L.KERNEL_PCG_B_FP64, 8)
    dom, dom_bytes, dom_ms, dom_n = dom64
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    log(f"FP64 Jacobi-PCG: {r64.iters} it, relres {r64.relres:.3e}, {t64:.3f} s/solve; dominant {dom}: "
        f"{dom_ms * 1e3:.1f} us avg over {dom_n} launches, {achieved:.0f} GB/s")
    # the mixed-precision pattern on the same system
    solve(cv.PRECISION_MIXED)
    xm, rm, t_mixed, domm, gbsm = timed(cv.PRECISION_MIXED, 2, L.KERNEL_PCG_A_FP32, L.KERNEL_PCG_B_FP32, 4)
    # error of the mixed solution against the FP64 one (mean-free: the operator
    # is singular up to its 1e-12*dvol shift)
    d = (xm - xm.mean()) - (x64 - x64.mean())
    err_vs_fp64 = (d.abs().max() / (x64 - x64.mean()).abs().max()).item()
    log(f"mixed: {rm.iters} it, relres {rm.relres:.3e}, {t_mixed:.3f} s/solve, dominant {domm[0]} "
        f"{domm[2] * 1e3:.1f} us")
    mixed = {"s": t_mixed, "iters": rm.iters, "relres": rm.relres, "err_vs_fp64": err_vs_fp64,
             "speedup_vs_fp64": t64 / t_mixed, "pcg_iteration_GBps": gbsm,
             "dominant_kernel": {"kernel": domm[0], "avg_launch_us": domm[2] * 1e3, "launches": domm[3],
                                 "bytes_per_launch": domm[1],
                                 "achieved_GBps": domm[1] / (domm[2] * 1e-3) / 1e9,
                                 "traffic": TRAFFIC_2048.get(domm[0]) if n == 2048 else None}}

    # ------------------------------------------- the PCG kernels at 4096^2
    # 2048 iterations of the FP64 and of the FP32 (mixed) Jacobi loop on the
    # same workload at 4096^2, where the FP32 working set (537 MB) no longer
    # fits the 256 MB Infinity Cache: the kernels' HBM-resident rates
    n4 = 2 * n
    case4 = wl.poisson_case(n4, seed=0)
    ctx4 = cv.Context(n4, n4, bc=case4["bc"])
    T4 = {k: torch.tensor(case4[k], device=dev) for k in ("u", "v", "rho_u", "rho_v")}
    b4 = ctx4.poisson_assemble(T4["u"], T4["v"], T4["rho_u"], T4["rho_v"], case4["dt"])
    del T4, case4
    p4 = {"workload": f"the same Poisson workload at {n4}x{n4}, 2048 Jacobi-PCG iterations (FP64, then the "
                      "mixed solve's FP32 recurrence); kernel times from HIP events on every 16th iteration"}
    for name, prec, tags, word in (("fp64", cv.PRECISION_FP64, (L.KERNEL_PCG_A_FP64, L.KERNEL_PCG_B_FP64), 8),
                                   ("fp32", cv.PRECISION_MIXED, (L.KERNEL_PCG_A_FP32, L.KERNEL_PCG_B_FP32), 4)):
        ctx4.poisson_solve(b4, ctx4.new_work(), precond=cv.PRECOND_JACOBI, tol=1e-10, max_iter=256, precision=prec)
        ctx4.profile(16)
        ctx4.poisson_solve(b4, ctx4.new_work(), precond=cv.PRECOND_JACOBI, tol=1e-10, max_iter=2048, precision=prec)
        (ms_a, n_a), (ms_b, n_b) = ctx4.profile_read(tags[0]), ctx4.profile_read(tags[1])
        ctx4.profile(0)
        ua, ub = ms_a / max(n_a, 1) * 1e3, ms_b / max(n_b, 1) * 1e3
        ba, bb = 8 * word * n4 * n4, 4 * word * n4 * n4
        p4[name] = {"pcg_a_us": ua, "pcg_a_GBps": ba / (ua * 1e-6) / 1e9, "pcg_a_frac": ba / (ua * 1e-6) / 1e9 /
                    HBM_PEAK_GBS, "pcg_b_us": ub, "pcg_b_GBps": bb / (ub * 1e-6) / 1e9, "launches": n_a}
        log(f"PCG kernels {n4}^2 {name}: A {ua:.1f} us ({p4[name]['pcg_a_GBps']:.0f} GB/s), B {ub:.1f} us")
    del ctx4, b4

    # ------------------------------------------- multigrid-preconditioned CG
    # the same system with the library's V-cycle preconditioner (no reference
    # counterpart), timed the same way: FP64; mixed (the default form: FP64 CG
    # whose V-cycle runs in FP32); and the FP32 Krylov recurrence with FP64
    # reliable updates (the form of the Jacobi mixed solve)
    mg = {}
    for name, prec, form in (("fp64", cv.PRECISION_FP64, 0), ("mixed", cv.PRECISION_MIXED, 0),
                             ("mixed_fp32_krylov", cv.PRECISION_MIXED, 1)):
        ctx.set_solver_option(cv.OPT_MIXED_MG_FORM, form)
        x = ctx.new_work()
        ctx.poisson_solve(b, x, precond=cv.PRECOND_MG, tol=tol, max_iter=20000, precision=prec)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 5
        for _ in range(reps):
            x = ctx.new_work()
            rmg = ctx.poisson_solve(b, x, precond=cv.PRECOND_MG, tol=tol, max_iter=20000, precision=prec)
        torch.cuda.synchronize()
        dt_mg = (time.perf_counter() - t) / reps
        assert rmg.converged and rmg.relres <= tol, rmg
        mg[name] = {"s": dt_mg, "iters": rmg.iters, "relres": rmg.relres}
        log(f"MG-PCG {name}: {rmg.iters} it, {dt_mg * 1e3:.1f} ms")
    ctx.set_solver_option(cv.OPT_MIXED_MG_FORM, 0)
    mg["speedup_mixed_mg_vs_fp64_mg"] = mg["fp64"]["s"] / mg["mixed"]["s"]
    mg["speedup_mixed_mg_vs_fp64_jacobi"] = t64 / mg["mixed"]["s"]
    mg["speedup_fp64_mg_vs_fp64_jacobi"] = t64 / mg["fp64"]["s"]

    # -------------------------------------------------------------------- VOF
    vn = n
    t0 = time.time()
    Cv, uv, vv, dtv = wl.vof_case(vn, seed=0)
    log(f"vof case {vn}^2 built in {time.time() - t0:.1f}s")
    vctx = cv.Context(vn, vn, bc=(cv.BC_PERIODIC,) * 4)
    Cg, ug, vg = (torch.tensor(a, device=dev) for a in (Cv, uv, vv))
    v0 = Cg[1:vn + 1, 1:vn + 1].sum().item()
    vctx.advect_vof(Cg, ug, vg, dtv, reconstruction_method="ELVIRA", nsteps=2)
    torch.cuda.synchronize()
    t = time.perf_counter()
    vctx.advect_vof(Cg, ug, vg, dtv, reconstruction_method="ELVIRA", nsteps=args.vof_steps)
    torch.cuda.synchronize()
    t_vof = (time.perf_counter() - t) / args.vof_steps
    v1 = Cg[1:vn + 1, 1:vn + 1].sum().item()
    # per-kernel durations from a separate run with every launch bracketed by
    # dispatch-timestamp events (kept out of the timed steps above)
    vctx.profile(1)
    vctx.advect_vof(Cg, ug, vg, dtv, reconstruction_method="ELVIRA", nsteps=4)
    rec = vctx.profile_read(L.KERNEL_VOF_RECON)
    upd = vctx.profile_read(L.KERNEL_VOF_UPDATE)
    vctx.profile(0)
    vof = {
        "workload": f"ELVIRA + Weymouth-Yue, {vn}x{vn} periodic, 128 seeded ellipses (exact cell fractions), "
                    f"CFL 0.25, {args.vof_steps} steps",
        "cell_updates_per_s": vn * vn / t_vof,
        "ms_per_step": t_vof * 1e3,
        "algorithmic_GBps": 50.0 * vn * vn / t_vof / 1e9,
        "recon_us_per_pass": rec[0] * 1e3 / max(rec[1], 1),
        "update_us_per_pass": upd[0] * 1e3 / max(upd[1], 1),
        "update_GBps": 25.0 * vn * vn / (upd[0] * 1e-3 / max(upd[1], 1)) / 1e9,
        "volume_drift_rel": abs(v1 - v0) / v0,
        "mixed_cell_fraction": ((Cg[1:vn + 1, 1:vn + 1] > 1e-6) & (Cg[1:vn + 1, 1:vn + 1] < 1 - 1e-6)).double()
        .mean().item(),
    }
    log(f"VOF: {t_vof * 1e3:.3f} ms/step, {vof['cell_updates_per_s']:.3e} cell-updates/s")
    # the same step at 4096^2 (the VOF workload's own size): 100 steps
    if vn == 2048 and not args.kernel_sample:
        del Cg, ug, vg
        vbig = 4096
        t0 = time.time()
        Cb, ub, vb, dtb = wl.vof_case(vbig, seed=0)
        log(f"vof case {vbig}^2 built in {time.time() - t0:.1f}s")
        bctx = cv.Context(vbig, vbig, bc=(cv.BC_PERIODIC,) * 4)
        Cg, ug, vg = (torch.tensor(a, device=dev) for a in (Cb, ub, vb))
        del Cb, ub, vb
        w0 = Cg[1:vbig + 1, 1:vbig + 1].sum().item()
        bctx.advect_vof(Cg, ug, vg, dtb, reconstruction_method="ELVIRA", nsteps=2)
        torch.cuda.synchronize()
        t = time.perf_counter()
        bctx.advect_vof(Cg, ug, vg, dtb, reconstruction_method="ELVIRA", nsteps=100)
        torch.cuda.synchronize()
        t_big = (time.perf_counter() - t) / 100
        w1 = Cg[1:vbig + 1, 1:vbig + 1].sum().item()
        vof["n4096"] = {"workload": "ELVIRA + Weymouth-Yue, 4096x4096 periodic, 128 seeded ellipses (exact cell "
                                    "fractions), CFL 0.25, 100 steps",
                        "ms_per_step": t_big * 1e3, "cell_updates_per_s": vbig * vbig / t_big,
                        "algorithmic_GBps": 50.0 * vbig * vbig / t_big / 1e9, "volume_drift_rel": abs(w1 - w0) / w0}
        log(f"VOF 4096^2: {t_big * 1e3:.3f} ms/step, {vbig * vbig / t_big:.3e} cell-updates/s")
        del bctx, Cg, ug, vg
        Cg = torch.tensor(Cv, device=dev)
    # the tail of advect_vof when surface tension is on: 3 diffusion-smoothing
    # steps (16 B/cell each + the halo pass) and curvature + CSF (3 passes:
    # normals 24 B/cell, curvature 24 B/cell, forces 32 B/cell)
    Ct = vctx.smooth_diffusion(Cg, 3)
    out_csf = vctx.compute_curvature_and_csf(Cg, 0.07, G=Ct)
    reps = 20
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        vctx.smooth_diffusion(Cg, 3, C_tilde=Ct)
    torch.cuda.synchronize()
    t_dif = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        vctx.compute_curvature_and_csf(Cg, 0.07, G=Ct, out=out_csf)
    torch.cuda.synchronize()
    t_csf = (time.perf_counter() - t) / reps
    vof["diffusion_smoothing_3_steps_us"] = t_dif * 1e6
    vof["curvature_csf_us"] = t_csf * 1e6
    vof["curvature_csf_GBps"] = 80.0 * vn * vn / t_csf / 1e9
    log(f"surface steps: diffusion x3 {t_dif * 1e6:.1f} us, curvature + CSF {t_csf * 1e6:.1f} us")

    # ------------------------------------------------- SP velocity predictor
    # the momentum block of the standard projection at the same size:
    # two-phase properties (rho 1000 / 100, mu 10 / 1 from seeded circles),
    # BDF2 with an n-1 state, TVD superbee convection, gravity, no-slip box,
    # CFL 0.25; BiCGSTAB(2) + Jacobi to 1e-10
    t0 = time.time()
    fm, dtm, bcm, lxm, lym = wl.momentum_case(n, n, seed=0)
    log(f"momentum case {n}^2 built in {time.time() - t0:.1f}s")
    mctx = cv.Context(n, n, lxm, lym, bc=bcm)
    mcfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98)
    base = {k: torch.tensor(fm[k], device=dev) for k in cv.Context.MOMENTUM_FIELDS}
    del fm
    n_unk = mctx.momentum_size()[0]
    reps = 3
    copies = [{**base, "u": base["u"].clone(), "v": base["v"].clone()} for _ in range(reps + 1)]
    mctx.solve_standard_projection_predictor(mcfg, copies[0], dtm, cv.LINEAR_BICGSTAB2, cv.PRECOND_JACOBI, tol,
                                             1000)  # warm-up
    mctx.profile(1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for k in range(reps):
        rp = mctx.solve_standard_projection_predictor(mcfg, copies[k + 1], dtm, cv.LINEAR_BICGSTAB2,
                                                      cv.PRECOND_JACOBI, tol, 1000)
    torch.cuda.synchronize()
    t_pred = (time.perf_counter() - t) / reps
    mv = mctx.profile_read(L.KERNEL_MOM_MATVEC)
    mctx.profile(0)
    assert rp.converged, rp
    # algorithmic bytes of one operator product per unknown: x, the two
    # velocity components of fields_old, rho, mu, mu_node, y (7 FP64 words)
    mv_us = mv[0] * 1e3 / max(mv[1], 1)
    predictor = {
        "workload": f"SP momentum block {n}x{n}: rho 1000/100, mu 10/1 (16 seeded circles), BDF2, TVD superbee, "
                    "gravity, no-slip walls, CFL 0.25; BiCGSTAB(2) + Jacobi to 1e-10",
        "unknowns": n_unk, "iters": rp.iters, "relres": rp.relres, "s": t_pred,
        "matvec_us": mv_us, "matvec_launches": mv[1],
        "matvec_GBps": 56.0 * n_unk / (mv_us * 1e-6) / 1e9,
    }
    log(f"predictor: {rp.iters} it, {t_pred * 1e3:.2f} ms, operator {mv_us:.1f} us "
        f"({predictor['matvec_GBps']:.0f} GB/s)")
    del copies, base, mctx

    # ------------------------------------------------------- the whole time step
    # the standard-projection time step of a two-phase run, all on the GPU
    # (transport, properties, CSF, velocity BC, predictor, Poisson,
    # correction): 4096 x 2048 on a 2 x 1 box, 256 seeded non-overlapping
    # bubbles (radii 8-32 cells, exact volume fractions), rho 1000 / 1,
    # mu 0.1 / 0.001, sigma 0.07, g 9.81, BDF2, TVD superbee, PLIC + WY;
    # predictor BiCGSTAB(2) + Jacobi to 1e-10 (the operator is not symmetric:
    # CG stalls once convection grows, at step 340 of this run); Poisson:
    # mixed-precision CG with the
    # multigrid preconditioner to 1e-10.  The CSF curvature from heights of C
    # (the full coupled step's specification), and the reference's -div(n)
    # beside it.
    sx, sy = 4096, 2048
    t0 = time.time()
    Cs, nbub = wl.bubbles_case(sx, sy, 256, 8.0, 32.0, seed=0, lx=2.0, ly=1.0, exact=True)
    log(f"bubbles case {sx}x{sy} ({nbub} bubbles) built in {time.time() - t0:.1f}s")
    sctx = cv.Context(sx, sy, 2.0, 1.0, bc=(cv.BC_FREESLIP, cv.BC_FREESLIP, cv.BC_DIRICHLET, cv.BC_DIRICHLET))
    zs = torch.zeros((sy + 2, sx + 2), dtype=torch.float64, device=dev)
    C_dev = torch.tensor(Cs, device=dev)
    del Cs
    steps_out = {}

    def make_run(curv, prec):
        # The mean-free coarsest solve of the V-cycle (the default for a
        # singular operator) takes this system's FP64 MG-PCG from about 250
        # to 150 iterations with the reference's curvature, but from 255 to
        # 285 with height-function curvature, whose solves end below the
        # FP64 floor of the true residual (|b| is 8x smaller); the
        # height-function runs keep the plain coarsest sweeps.
        sctx.set_solver_option(cv.OPT_MG_PROJECT, 0 if curv == "height_function" else 1)
        return cv.StandardProjectionRun(sctx, C_dev, zs, zs, 1000.0, 1.0, 0.1, 1.0e-3, 0.07,
                                        time_scheme=cv.TIME_SCHEME_GEAR, convective_scheme=cv.CONV_TVD_SUPERBEE,
                                        gy=-9.81, predictor=(cv.LINEAR_BICGSTAB2, cv.PRECOND_JACOBI, 1e-10, 500),
                                        poisson=(cv.LINEAR_CG, cv.PRECOND_MG, 1e-10, 20000, prec), curvature=curv)

    def describe(curv, prec, nst):
        return (f"two-phase standard-projection step {sx}x{sy} (2 x 1 box), {nbub} seeded bubbles (radii 8-32 "
                f"cells, exact volume fractions), rho 1000/1, mu 0.1/0.001, sigma 0.07 (CSF, curvature "
                f"{curv.replace('_', '-')}), g 9.81, BDF2, TVD superbee, PLIC + WY, dt 1e-4, {nst} steps; "
                "predictor BiCGSTAB(2) + Jacobi to 1e-10; Poisson " + ("mixed-precision (FP64 CG, FP32 V-cycle)"
                                                             if prec == cv.PRECISION_MIXED
                                                             else "FP64") + " MG-PCG to 1e-10")

    dts = 1.0e-4
    # the full coupled step's own length: 500 steps with height-function
    # curvature and the mixed-precision Poisson solve (after two warm-up steps
    # of another run on the same context: graph captures, code objects)
    warm = make_run("height_function", cv.PRECISION_MIXED)
    for _ in range(2):
        warm.step(dts)
    del warm
    run = make_run("height_function", cv.PRECISION_MIXED)
    vol0 = run.f["C"][1:sy + 1, 1:sx + 1].sum().item()
    n500 = args.sp_steps
    torch.cuda.synchronize()
    t = time.perf_counter()
    its, umax = [], 0.0
    for k in range(n500):
        its.append(run.step(dts))
        if (k + 1) % 25 == 0:
            umax = max(umax, run.f["u"].abs().max().item(), run.f["v"].abs().max().item())
        if (k + 1) % 100 == 0:
            log(f"  SP step {k + 1}/{n500}")
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t) / n500
    vol1 = run.f["C"][1:sy + 1, 1:sx + 1].sum().item()
    pit = [a[1] for a in its]
    sp_step = {"workload": describe("height_function", cv.PRECISION_MIXED, n500), "steps": n500,
               "ms_per_step": t_step * 1e3, "cell_steps_per_s": sx * sy / t_step,
               "predictor_iters_mean": sum(a[0] for a in its) / n500, "poisson_iters_mean": sum(pit) / n500,
               "poisson_iters_max": max(pit), "all_converged": all(a[2] == 0 and a[3] == 0 for a in its),
               "volume_drift_rel": abs(vol1 - vol0) / vol0, "kinetic_energy": run.kinetic_energy(),
               "max_cfl_sampled": umax * dts / (2.0 / sx)}
    log(f"SP step {sx}x{sy} x{n500}: {t_step * 1e3:.2f} ms/step, Poisson {sp_step['poisson_iters_mean']:.0f} it "
        f"(max {max(pit)}), all converged {sp_step['all_converged']}, KE {sp_step['kinetic_energy']:.3e}")
    Cw, uw, vw = (run.f[k].clone() for k in ("C", "u", "v"))
    del run
    # the reference's curvature and the FP64 Poisson solve beside it, 5 steps each
    for curv, prec in (("normals", cv.PRECISION_MIXED), ("height_function", cv.PRECISION_FP64)):
        run = make_run(curv, prec)
        for _ in range(2):
            run.step(dts)
        torch.cuda.synchronize()
        t = time.perf_counter()
        nst = 5
        sp_its = [run.step(dts) for _ in range(nst)]
        torch.cuda.synchronize()
        t_v = (time.perf_counter() - t) / nst
        steps_out[(curv, prec)] = {
            "workload": describe(curv, prec, nst), "ms_per_step": t_v * 1e3, "cell_steps_per_s": sx * sy / t_v,
            "predictor_iters": [a[0] for a in sp_its], "poisson_iters": [a[1] for a in sp_its],
            "status": [[a[2], a[3]] for a in sp_its], "kinetic_energy": run.kinetic_energy()}
        log(f"SP step {sx}x{sy} ({curv}, precision {prec}): {t_v * 1e3:.2f} ms/step, Poisson "
            f"{steps_out[(curv, prec)]['poisson_iters']} it")
        del run
    sp_step["reference_curvature"] = steps_out[("normals", cv.PRECISION_MIXED)]
    sp_step["fp64_poisson"] = steps_out[("height_function", cv.PRECISION_FP64)]

    # ------------------------- the 16-step checkpointed kinetic-energy adjoint
    # dJ/dC0 of J = KE after 16 steps, from the state the 500-step run ended
    # in, with the momentum block reduced to BDF2 inertia and gravity
    # (KineticEnergyAdjoint): forward with per-step checkpoints, then the
    # reverse sweep; forward and transpose solves FP64 MG-PCG to 1e-12; one
    # central difference (h = 1e-5) along a seeded direction on the mixed
    # cells as a check (entries uniform in [0, 1]: the gradient there is
    # mostly positive, so the directional derivative stands well above the
    # ~1e-13 relative rounding noise of J, which a zero-mean direction's
    # near-cancelling derivative would not)
    nwin = 16
    sctx.set_solver_option(cv.OPT_MG_PROJECT, 0)  # the height-function state of the 500-step run (see make_run)
    kadj = cv.KineticEnergyAdjoint(sctx, 1000.0, 1.0, dts, gy=-9.81, reconstruction="PLIC", tol=1e-12,
                                   precond=cv.PRECOND_MG)
    kadj.forward(Cw, uw, vw, 1)  # warm-up: graph captures, work arrays
    kadj.backward()
    torch.cuda.synchronize()
    t = time.perf_counter()
    Jk = kadj.forward(Cw, uw, vw, nwin)
    torch.cuda.synchronize()
    t_kf = time.perf_counter() - t
    t = time.perf_counter()
    gC, _, _ = kadj.backward()
    torch.cuda.synchronize()
    t_kb = time.perf_counter() - t
    fwd_iters = list(kadj.iters)
    Cin = Cw[1:sy + 1, 1:sx + 1]
    mixed_cells = (Cin >= 0.02) & (Cin <= 0.98)
    gen = torch.Generator(device=dev).manual_seed(11)
    dk = torch.zeros_like(Cw)
    dk[1:sy + 1, 1:sx + 1] = torch.rand(Cin.shape, generator=gen, device=dev, dtype=torch.float64) * mixed_cells
    adk = float((gC * dk).sum())
    gscale = float(gC[1:sy + 1, 1:sx + 1][mixed_cells].norm())
    hk = 1e-5
    fdk = (kadj.forward(Cw + hk * dk, uw, vw, nwin) - kadj.forward(Cw - hk * dk, uw, vw, nwin)) / (2.0 * hk)
    nh_w = (sx + 2) * (sy + 2)
    ke_adjoint = {"workload": f"16-step checkpointed adjoint of the kinetic energy w.r.t. C0, {sx}x{sy} walls, "
                              f"from the state after the {n500}-step run ({nbub} bubbles), rho 1000/1, g 9.81, "
                              "dt 1e-4, PLIC + WY, momentum block reduced to BDF2 inertia + gravity; forward and "
                              "transpose FP64 MG-PCG to 1e-12",
                  "steps": nwin, "kinetic_energy": Jk, "forward_s": t_kf, "backward_s": t_kb,
                  "forward_poisson_iters": fwd_iters, "checkpoint_bytes": 4 * 8 * nh_w * (nwin + 1),
                  "fd_direction": {"h": hk, "fd": fdk, "ad": adk, "grad_norm_mixed": gscale,
                                   "rel_err": abs(fdk - adk) / abs(adk)}}
    log(f"KE adjoint {sx}x{sy} x{nwin}: forward {t_kf:.2f} s, backward {t_kb:.2f} s, FD rel "
        f"{ke_adjoint['fd_direction']['rel_err']:.2e}")
    del kadj, gC, dk, Cw, uw, vw
    del sctx, zs, C_dev

    # --------------------------------------------- the adjoint of the projection
    # 1024^2, density of seeded circles (rho 1 vs 1000, one-cell smoothing),
    # J = sum(w p) with w seeded Gaussian: the forward solve and the transpose
    # solve of dJ/d(log rho) and dJ/d(rhs) through torch.autograd, both in the
    # mixed mode (FP32 inner, FP64 refined) to 1e-12, checked against a
    # fourth-order central difference along one seeded direction
    from covof_amd.autograd import pressure_solve
    na = 1024
    acase = wl.poisson_case(na, seed=0, n_circles=64, rmin=8, rmax=64)
    Ca = acase["C"][1:na + 1, 1:na + 1]
    log_rho = np.log(1000.0 * Ca + (1.0 - Ca))
    rng_a = np.random.default_rng(7)
    w_a = rng_a.normal(size=(na, na))
    w_a -= w_a.mean()
    actx = cv.Context(na, na, bc=acase["bc"])
    wg = torch.tensor(w_a, device=dev)
    bg = torch.tensor(acase["rhs"], device=dev)
    aopts = dict(tol=1e-12, max_iter=400000, precision=cv.PRECISION_MIXED)
    lr = torch.tensor(log_rho, device=dev).requires_grad_(True)
    torch.cuda.synchronize()
    t = time.perf_counter()
    Jv = (wg * pressure_solve(actx, lr, bg, 1.0, **aopts)).sum()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t
    t = time.perf_counter()
    Jv.backward()
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t
    ga = lr.grad.cpu().numpy()

    def J_of(lrv):
        with torch.no_grad():
            return (wg * pressure_solve(actx, torch.tensor(lrv, device=dev), bg, 1.0, **aopts)).sum().item()

    ha = 1e-3
    da = rng_a.normal(size=log_rho.shape)
    d1 = J_of(log_rho + ha * da) - J_of(log_rho - ha * da)
    d2 = J_of(log_rho + 2 * ha * da) - J_of(log_rho - 2 * ha * da)
    fd = (8.0 * d1 - d2) / (12.0 * ha)
    ad = float((ga * da).sum())
    adjoint = {"workload": f"adjoint of the projection {na}x{na}: rho 1 vs 1000 (64 seeded circles, radii 8-64 "
                           "cells, one-cell smoothing), J = sum(w p), w seeded Gaussian; forward and transpose "
                           "Jacobi-PCG solves in the mixed mode to 1e-12",
               "forward_s": t_fwd, "adjoint_s": t_bwd, "fd_rel_err": abs(fd - ad) / abs(ad)}
    log(f"adjoint {na}^2: forward {t_fwd:.3f} s, adjoint {t_bwd:.3f} s, vs FD {adjoint['fd_rel_err']:.2e}")
    del actx, lr, wg, bg

    # ------------------------------------------ the plain-method case (128 x 256)
    # the reference's own CPU-runnable case: 2-D rising bubble on a 1 x 2 box,
    # rho 1000 / 100, mu 10 / 1, sigma 24.5, g 0.98, circle r 0.25 at (0.5, 0.5)
    # (set_initial_conditions' BUBBLE), BDF2, TVD superbee, PLIC + WY, dt 1e-3,
    # 200 steps; predictor CG + Jacobi to 1e-12, the reference's hard-coded
    # Poisson block (CG + exact symmetric Gauss-Seidel to 1e-7)
    import tempfile
    from covof_amd import case as cvcase
    pj = {"case_name": "plain", "mode": "NAVIER_STOKES", "mesh": {"nx": 128, "ny": 256, "lx": 1.0, "ly": 2.0},
          "physics": {"enable_two_phase": True, "enable_surface_tension": True, "sigma": 24.5, "gy": -0.98,
                      "convective_scheme": "TVD_SUPERBEE", "fluid1": {"rho": 1000.0, "mu": 10.0},
                      "fluid2": {"rho": 100.0, "mu": 1.0}},
          "time": {"time_scheme": "BDF2", "dt": 1e-3, "t_final": 0.2},
          "solver": {"type": "STANDARD_PROJECTION", "linear_solver": "CG", "preconditioner": "JACOBI",
                     "tol": 1e-12, "max_iter": 500},
          "vof": {"advection_method": "WEYMOUTH_YUE", "reconstruction_method": "PLIC"},
          "boundary_conditions": {"left": {"type": "FREE_SLIP"}, "right": {"type": "FREE_SLIP"},
                                  "bottom": {"type": "DIRICHLET"}, "top": {"type": "DIRICHLET"}},
          "initial_condition": {"type": "BUBBLE", "center_x": 0.5, "center_y": 0.5, "radius": 0.25}}
    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as fj:
        json.dump(pj, fj)
    pc = cvcase.load_config(fj.name)
    os.unlink(fj.name)
    C0 = cvcase.initial_volume_fraction(pc)
    cvcase._vof_halo(C0, pc)
    pctx = cv.Context(128, 256, 1.0, 2.0, bc=(cv.BC_FREESLIP, cv.BC_FREESLIP, cv.BC_DIRICHLET, cv.BC_DIRICHLET))
    z0 = torch.zeros((258, 130), dtype=torch.float64, device=dev)
    prun = cv.StandardProjectionRun(pctx, torch.tensor(C0, device=dev), z0, z0, 1000.0, 100.0, 10.0, 1.0, 24.5,
                                    time_scheme=cv.TIME_SCHEME_GEAR, convective_scheme=cv.CONV_TVD_SUPERBEE,
                                    gy=-0.98, predictor=(cv.LINEAR_CG, cv.PRECOND_JACOBI, 1e-12, 500))
    f0 = prun.f
    pctx.compute_curvature_and_csf(f0["C"], 24.5, out={k: f0[k] for k in ("curvature", "fx_sigma", "fy_sigma")})
    pctx.apply_velocity_bc(prun.cfg(False), f0["u"], f0["v"])
    torch.cuda.synchronize()
    t = time.perf_counter()
    p_its = [prun.step(1e-3) for _ in range(200)]
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t) / 200
    plain = {"workload": "plain-method rising bubble 128x256, 200 steps of dt 1e-3 (the reference's configuration; "
                         "Poisson CG + exact SGS to 1e-7)",
             "ms_per_step": t_plain * 1e3, "predictor_iters_total": sum(a[0] for a in p_its),
             "poisson_iters_total": sum(a[1] for a in p_its), "kinetic_energy": prun.kinetic_energy()}
    log(f"plain case: {t_plain * 1e3:.2f} ms/step, iterations {plain['predictor_iters_total']} + "
        f"{plain['poisson_iters_total']}")
    del prun, pctx

    # ---------------------------------------------- the monolithic (coupled) step
    # one coupled solve of u, v, p (one Uzawa pass) on the momentum case at
    # 128^2 and 256^2: BiCGSTAB(2) + block-triangular Schur-SSOR to 1e-10
    mono = {"workload": "monolithic step on the momentum case (rho 1000/100, mu 10/1, BDF2, TVD, no-slip box): "
                        "BiCGSTAB(2) + block-triangular Schur-SSOR to 1e-10"}
    for mn in (128, 256):
        fmn, dtmn, bcmn, lxmn, lymn = wl.momentum_case(mn, mn, seed=0)
        mctx = cv.Context(mn, mn, lxmn, lymn, bc=bcmn)
        mcfg = cv.Context.momentum_cfg(time_scheme=2, has_nm1=True, convective_scheme=3, gy=-0.98,
                                       solver_type=cv.SOLVER_MONOLITHIC)
        mbase = {k: torch.tensor(fmn[k], device=dev) for k in cv.Context.MOMENTUM_FIELDS}
        mctx.momentum_set_fields(mcfg, mbase, dtmn)
        rvec = mctx.new_momentum_vector(1.0)
        mctx.apply_preconditioner(rvec, precond=cv.PRECOND_BLOCK_TRIANGULAR)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(3):
            mctx.apply_preconditioner(rvec, precond=cv.PRECOND_BLOCK_TRIANGULAR)
        torch.cuda.synchronize()
        t_bt = (time.perf_counter() - t) / 3
        pm = torch.zeros((mn + 2, mn + 2), dtype=torch.float64, device=dev)
        t = time.perf_counter()
        rmn = mctx.solve_monolithic_step(mcfg, mbase, pm, dtmn, cv.LINEAR_BICGSTAB2, cv.PRECOND_BLOCK_TRIANGULAR,
                                         1e-10, 2000)
        torch.cuda.synchronize()
        t_mono = time.perf_counter() - t
        mono[f"n{mn}"] = {"unknowns": mctx.momentum_size()[0], "iters": rmn.iters, "relres": rmn.relres,
                          "status": rmn.status, "s": t_mono, "block_triangular_ms": t_bt * 1e3}
        log(f"monolithic {mn}^2: {rmn.iters} it, {t_mono:.3f} s, block-triangular {t_bt * 1e3:.2f} ms")
        del mctx, mbase, rvec, pm

    # ----------------------------------------------------------- CPU baseline
    # the oracle (the C restatement of the reference's algorithm) on the host
    # cores: one whole FP64 Jacobi-PCG solve of the same 2048^2 system to 1e-10
    # (OpenMP), and a few ELVIRA + WY steps of the 2048^2 VOF workload (the
    # oracle's cell loops on the same threads)
    cpu, cpu_vof = None, None
    if not args.no_cpu:
        from oracle import oracle as orc
        threads = min(16, os.cpu_count() or 1)
        m = orc.mesh(n, n, bc=case["bc"])
        co = orc.poisson_assemble(m, case["u"], case["v"], case["rho_u"], case["rho_v"], case["dt"])
        sec, cit, crr = orc.bench_jacobi_pcg_solve(m, co, co["b"], tol, 400000, threads)
        cpu = {"value": sec, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"one whole FP64 Jacobi-PCG solve of the {n}^2 system to 1e-10 on the host "
                         f"({cit} iterations, relres {crr:.2e}; oracle, OpenMP {threads} threads, "
                         f"{sec / cit * 1e3:.2f} ms/iteration)"}
        log(f"CPU Poisson: {sec:.2f} s, {cit} it on {threads} threads")
        del co
        orc.set_threads(threads)
        mv_ = orc.mesh(vn, vn, bc=(3, 3, 3, 3))
        Co, sc = Cv.copy(), 0
        Co, sc = orc.wy_step(mv_, orc.ELVIRA, Co, uv, vv, dtv, sc)  # warm-up
        t = time.perf_counter()
        for _ in range(args.cpu_vof_steps):
            Co, sc = orc.wy_step(mv_, orc.ELVIRA, Co, uv, vv, dtv, sc)
        t_cv = (time.perf_counter() - t) / args.cpu_vof_steps
        orc.set_threads(1)
        cpu_vof = {"value": vn * vn / t_cv, "unit": "cell-updates/s", "cores": threads, "kind": "port",
                   "ms_per_step": t_cv * 1e3,
                   "sample": f"{args.cpu_vof_steps} ELVIRA + Weymouth-Yue steps of the {vn}^2 VOF workload "
                             f"(oracle, OpenMP {threads} threads over the cell loops)"}
        log(f"CPU VOF: {t_cv * 1e3:.1f} ms/step on {threads} threads")

    out = {
        "metric": METRIC,
        "value": t64,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"variable-density Poisson {n}x{n}: 64 seeded circles (radii 16-128 cells), "
                               "rho 1 vs 1000 with 1-cell smoothing, zero-mean N(0,1) rhs (seed 0), walls; "
                               "FP64 Jacobi-PCG to 1e-10",
                   "n": n, "tol": tol, "precond": "JACOBI"},
        "iters": r64.iters,
        "relres": r64.relres,
        "pcg_iteration_GBps": gbs64,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_2048.get(dom) if n == 2048 else None,
                     "bytes_per_launch": dom_bytes, "avg_launch_us": dom_ms * 1e3, "launches": dom_n},
        "cpu_baseline": cpu,
        "cpu_baseline_vof": cpu_vof,
        "mixed_precision": mixed,
        "pcg_kernels_4096": p4,
        "multigrid_pcg": mg,
        "vof": vof,
        "sp_predictor": predictor,
        "sp_step": sp_step,
        "adjoint": adjoint,
        "ke_adjoint_16step": ke_adjoint,
        "plain_case": plain,
        "monolithic": mono,
    }
    print(json.dumps(_plain(out)), flush=True)


def _plain(o, path="out"):
    """The result with one-element tensors as numbers (and a log line naming
    each, so the producing expression can be fixed)."""
    import torch
    if isinstance(o, dict):
        return {k: _plain(v, f"{path}.{k}") for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_plain(v, f"{path}[{i}]") for i, v in enumerate(o)]
    if isinstance(o, torch.Tensor):
        log(f"note: tensor value at {path}")
        return o.item() if o.numel() == 1 else o.tolist()
    return o


if __name__ == "__main__":
    main()
