"""Benchmark: quasi-dynamic earthquake-cycle steps per second at 1001^2.

Workload (the headline configuration): 2-D antiplane strike-slip QD cycle,
Ny = Nz = 1001, Ly = Lz = 30 km, 4th-order SBP-SAT, sinh-stretched z grid
(stretch 5), layered shear modulus 30 -> 36 GPa over the top 4 km, aging-law
rate-and-state friction (a ramps 0.015 -> 0.025 between 12 and 17 km),
adaptive RK32 with PID control (timeStepTol 1e-8), synthetic inputs with a
seeded +-1% state perturbation.  A "step" is one accepted RK32 step (three
d_dt evaluations, each a full elasticity solve to relative residual 1e-12,
plus rejected attempts if any).

Two GPU paths run the same steps from the same state: the mixed-precision
solve (FP32 Jacobi-PCG inside FP64 iterative refinement) -- `value` -- and the
plain FP64 Jacobi-PCG.  The line also carries the slip relative error of the
mixed path against the FP64 path, the HBM roofline fraction of the dominant
kernel (the FP32 interior SBP apply, timed with HIP events on its stream) and
a CPU baseline (the FP64 oracle's Jacobi-PCG on the assembled operator, timed
on a bounded sample and scaled by the iterations per step of the FP64 run).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PATHS = {
    # label: (solve mode, preconditioner)
    "mixed": ("cg64", "fdiag32"),      # FP64 CG, FP32 fast-diagonalisation preconditioner
    "mixed_bf16": ("cg64", "fdiag_bf16"),
    "fp64_fdiag": ("cg64", "fdiag64"),
    "fp64": ("cg64", "jacobi"),        # plain FP64 Jacobi-PCG
    "ir32": ("ir32", "jacobi"),        # FP32 Jacobi-PCG inside FP64 refinement
}


def run_path(cfg, label, warmup, steps, reuse=0):
    from scycle_amd import QdMediator
    mode, prec = PATHS[label]
    tic = time.perf_counter()
    q = QdMediator(cfg, solve=mode, rtol=1e-12, inner_rtol=1e-5, jac_mode=1, precond=prec,
                   reuse=reuse)
    setup = time.perf_counter() - tic
    q.set_state(cfg.psi0, cfg.slip0, np.zeros(cfg.Nz))
    t0 = 0.0
    dt0 = 1e-3
    if warmup > 0:
        t0, n, dts = q.integrate(warmup, integrator="RK32", tol=1e-8, init_dt=dt0, max_dt=1e10)
        dt0 = float(dts[-1])
    c0 = q.counters()
    tic = time.perf_counter()
    t1, n, dts = q.integrate(steps, integrator="RK32", tol=1e-8, init_dt=dt0, init_time=t0,
                             max_dt=1e10)
    sec = time.perf_counter() - tic
    c1 = q.counters()
    st = q.get_state()
    d = {k: c1[k] - c0[k] for k in ("d_dt_calls", "solves", "cg_iters", "outer_sweeps")}
    return q, dict(seconds=sec, steps=n, t_start=t0, t_end=t1, counters=d, state=st,
                   setup_s=setup, dt_min=float(dts.min()), dt_max=float(dts.max()))


def roofline(q):
    """FP32 interior apply kernel: 12 algorithmic bytes per node (read u, read
    mu, write A*u), timed with HIP events on the context stream."""
    from scycle_amd._lib import check, load
    ctx = load().sc_qd_ctx(q.h)
    ms, nodes = C.c_double(), C.c_double()
    check(load().sc_bench_apply(ctx, 3, 200, C.byref(ms), C.byref(nodes)))
    b = 12.0 * nodes.value
    ach = b / (ms.value * 1e-3)
    out = dict(kernel="k_apply_interior<float,4>", ms=ms.value, bytes=b, achieved=ach)
    check(load().sc_bench_apply(ctx, 2, 200, C.byref(ms), C.byref(nodes)))
    out["fp64_ms"] = ms.value
    out["fp64_GBps"] = 24.0 * nodes.value / (ms.value * 1e-3) / 1e9
    return out


def cpu_baseline(cfg, iters_per_step, budget_s=20.0):
    """FP64 oracle Jacobi-PCG at the same operator: time a bounded number of
    iterations, scale by the FP64 GPU run's CG iterations per step."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    tic = time.perf_counter()
    o = ol.OracleSbp(cfg.order, cfg.Ny, cfg.Nz, cfg.Ly, cfg.Lz, cfg.mu, cfg.y, cfg.z,
                     var_grid=cfg.var_grid, bc=cfg.mat_bc)
    t_build = time.perf_counter() - tic
    rng = np.random.default_rng(0)
    rhs = o.set_rhs(1e-3 * rng.random(cfg.Nz), np.full(cfg.Nz, 5e-4), np.zeros(cfg.Ny),
                    np.zeros(cfg.Ny))
    it = 20
    tic = time.perf_counter()
    _, done, _ = o.pcg(rhs, rtol=1e-300, maxit=it)
    t_probe = time.perf_counter() - tic
    per = t_probe / max(done, 1)
    it2 = int(max(10, min(2000, (budget_s - t_probe) / per)))
    tic = time.perf_counter()
    _, done2, _ = o.pcg(rhs, rtol=1e-300, maxit=it2)
    t_iter = (time.perf_counter() - tic) / max(done2, 1)
    step_s = iters_per_step * t_iter
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=1.0 / step_s if step_s > 0 else None, unit="QD steps/s", cores=cores,
                kind="port",
                sample=(f"{done2} Jacobi-PCG iterations of the FP64 oracle on the assembled "
                        f"{cfg.Ny}x{cfg.Nz} order-{cfg.order} operator ({t_iter*1e3:.1f} ms/iter, "
                        f"assembly {t_build:.1f} s untimed), scaled by "
                        f"{iters_per_step:.0f} CG iterations per RK32 step of the FP64 GPU run"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="mixed,fp64,fp64_fdiag,mixed_bf16")
    ap.add_argument("--reuse", type=int, default=0,
                    help="previous solutions projected for the CG initial guess (FD paths)")
    ap.add_argument("--slow-steps", type=int, default=3,
                    help="timed steps for the Jacobi-preconditioned paths (same window start)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus > 1 is not supported "
              "(the multi-GPU decomposition is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    from scycle_amd.configs import config_c2
    cfg = config_c2(args.n, args.n)
    res = {}
    keep = None
    labels = args.paths.split(",")
    for label in labels:
        k = args.steps if PATHS[label][1] != "jacobi" else min(args.steps, args.slow_steps)
        log(f"[bench] path {label} {PATHS[label]}: warmup {args.warmup} steps, timed {k} steps, "
            f"{cfg.Ny}x{cfg.Nz}")
        q, r = run_path(cfg, label, args.warmup, k,
                        reuse=args.reuse if PATHS[label][1] != "jacobi" else 0)
        log(f"[bench] {label}: {r['steps']} steps in {r['seconds']:.3f} s (setup {r['setup_s']:.1f} s), "
            f"counters {r['counters']}")
        res[label] = r
        if label == labels[0]:
            keep = q
        else:
            q.close()
    main_mode = labels[0]
    r = res[main_mode]
    value = r["steps"] / r["seconds"]
    ms_step = 1e3 * r["seconds"] / r["steps"]
    extra = {"paths": {}}
    for lab, rr in res.items():
        extra["paths"][lab] = dict(
            solve=PATHS[lab][0], precond=PATHS[lab][1], steps=rr["steps"],
            steps_per_s=rr["steps"] / rr["seconds"],
            cg_iters_per_solve=rr["counters"]["cg_iters"] / max(rr["counters"]["solves"], 1),
            solves_per_step=rr["counters"]["solves"] / rr["steps"], setup_s=rr["setup_s"],
            t_start=rr["t_start"], t_end=rr["t_end"], dt_min=rr["dt_min"], dt_max=rr["dt_max"])
    if "fp64" in res:
        extra["fp64_steps_per_s"] = extra["paths"]["fp64"]["steps_per_s"]
        extra["mixed_over_fp64"] = value / extra["fp64_steps_per_s"]
    if "fp64_fdiag" in res:
        extra["mixed_over_fp64_same_preconditioner"] = value / extra["paths"]["fp64_fdiag"]["steps_per_s"]
    ref = "fp64_fdiag" if "fp64_fdiag" in res else ("fp64" if "fp64" in res else None)
    if ref and ref != main_mode:
        # compare at the same step count: the reference path ran the same window
        n = min(res[ref]["steps"], r["steps"])
        if n == r["steps"] == res[ref]["steps"]:
            a, b = r["state"]["slip"], res[ref]["state"]["slip"]
            extra["slip_rel_err_vs_fp64"] = float(np.abs(a - b).max() / np.abs(b).max())
            extra["V_rel_err_vs_fp64"] = float(np.abs(r["state"]["V"] - res[ref]["state"]["V"]).max()
                                               / np.abs(res[ref]["state"]["V"]).max())
            extra["slip_err_reference_path"] = ref
    rl = roofline(keep)
    log(f"[bench] roofline {rl}")
    roof = dict(bound="hbm", achieved=rl["achieved"] / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=rl["achieved"] / HBM_PEAK, traffic=None, kernel=rl["kernel"],
                kernel_ms=rl["ms"], bytes_per_launch=rl["bytes"],
                fp64_interior_GBps=rl["fp64_GBps"])
    cpu = None
    if not args.no_cpu and "fp64" in res:
        ips = res["fp64"]["counters"]["cg_iters"] / res["fp64"]["steps"]
        log(f"[bench] cpu baseline (oracle), {ips:.0f} iterations/step")
        cpu = cpu_baseline(cfg, ips)
    line = {
        "metric": "QD steps/s at 1001^2 mixed vs FP64 GPU (slip rel.err <=1e-8) + HBM GB/s",
        "value": value, "unit": "QD steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f64 CG + f32 GEMM preconditioner", "mixed_bf16": "f64 CG + bf16 GEMM preconditioner",
                  "fp64_fdiag": "f64", "fp64": "f64", "ir32": "f32 CG + f64 refinement"}[main_mode],
        "data": "synthetic",
        "config": {"workload": f"QD strike-slip {cfg.Ny}x{cfg.Nz} order-4 sinh-z(5) layered-mu "
                               f"aging-law RK32 tol 1e-8, steps {args.warmup + 1}-"
                               f"{args.warmup + args.steps} from t=0",
                   "solve": main_mode, "rtol": 1e-12},
        "roofline": roof, "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
