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


def run_path(cfg, mode, warmup, steps):
    from scycle_amd import QdMediator
    q = QdMediator(cfg, solve=mode, rtol=1e-12, inner_rtol=1e-5, jac_mode=1)
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
    return q, dict(seconds=sec, steps=n, t_end=t1, counters=d, state=st)


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
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1001)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--paths", default="ir32,cg64")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus > 1 is not supported "
              "(the multi-GPU decomposition is a design note, not a measured path)",
              file=sys.stderr)
        sys.exit(2)
    from scycle_amd.configs import config_c2
    cfg = config_c2(args.n, args.n)
    res = {}
    q_mixed = None
    for mode in args.paths.split(","):
        log(f"[bench] path {mode}: warmup {args.warmup} steps, timed {args.steps} steps, "
            f"{cfg.Ny}x{cfg.Nz}")
        q, r = run_path(cfg, mode, args.warmup, args.steps)
        log(f"[bench] {mode}: {r['steps']} steps in {r['seconds']:.3f} s, counters {r['counters']}")
        res[mode] = r
        if mode == "ir32":
            q_mixed = q
        else:
            q.close()
    main_mode = "ir32" if "ir32" in res else list(res)[0]
    r = res[main_mode]
    value = r["steps"] / r["seconds"]
    ms_step = 1e3 * r["seconds"] / r["steps"]
    extra = {}
    if "ir32" in res and "cg64" in res:
        a, b = res["ir32"]["state"]["slip"], res["cg64"]["state"]["slip"]
        extra["fp64_steps_per_s"] = res["cg64"]["steps"] / res["cg64"]["seconds"]
        extra["mixed_over_fp64"] = value / extra["fp64_steps_per_s"]
        extra["slip_rel_err_mixed_vs_fp64"] = float(np.abs(a - b).max() / np.abs(b).max())
        extra["V_rel_err_mixed_vs_fp64"] = float(
            np.abs(res["ir32"]["state"]["V"] - res["cg64"]["state"]["V"]).max()
            / np.abs(res["cg64"]["state"]["V"]).max())
        extra["cg_iters_per_step"] = {m: res[m]["counters"]["cg_iters"] / res[m]["steps"]
                                      for m in res}
    rl = roofline(q_mixed if q_mixed is not None else q)
    log(f"[bench] roofline {rl}")
    roof = dict(bound="hbm", achieved=rl["achieved"] / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                frac=rl["achieved"] / HBM_PEAK, traffic=None, kernel=rl["kernel"],
                kernel_ms=rl["ms"], bytes_per_launch=rl["bytes"],
                fp64_interior_GBps=rl["fp64_GBps"])
    cpu = None
    if not args.no_cpu:
        ips = (res["cg64"]["counters"]["cg_iters"] / res["cg64"]["steps"]) if "cg64" in res else \
            r["counters"]["cg_iters"] / r["steps"]
        log(f"[bench] cpu baseline (oracle), {ips:.0f} iterations/step")
        cpu = cpu_baseline(cfg, ips)
    line = {
        "metric": "QD steps/s at 1001^2 mixed vs FP64 GPU (slip rel.err <=1e-8) + HBM GB/s",
        "value": value, "unit": "QD steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32-cg+fp64-ir" if main_mode == "ir32" else "f64",
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
