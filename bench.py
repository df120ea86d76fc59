"""Benchmark: particle-steps/s of the explicit SPH Leapfrog step on one MI355X.

Workload: cylinder upsetting, R=30 mm, L=60 mm, dx=0.5537 mm -> 976,320
particles, h=1.2 dx, +-1e-3 dx seeded jitter, Johnson-Cook hardening, top row
driven at -1 m/s (weldform_amd.cases.config2).  A step is one iteration of
SolveDiffUpdateLeapFrog's loop: neighbour schedule (rebuild when due),
acceleration gather, kicks + BCs, density + drift, strain/rotation rates +
stress update, energies and reaction reduction.  All inputs are resident in
HBM before the timed region.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  The cpu_baseline leg times the CPU oracle (OpenMP C
restatement of the reference path) on the same workload for a few steps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "particle-steps/s at 1M (FP32-mix vs FP64 GPU, stated err); adjoint grads/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kind, n, nbar):
    """Algorithmic HBM bytes per particle for one launch (FP64, SoA, each input
    read once, each output written once, int32 neighbour list read once)."""
    lst = 4.0 * nbar + 4.0
    if kind == "accel":          # x,h,v,m,rho,sigma in (120) ; a out (24)
        return n * (120.0 + 24.0 + lst)
    if kind == "density":        # x,h,v,m,rho in (72) ; rho,x,disp out (8+24+24) + m/rho (8)
        return n * (72.0 + 64.0 + lst)
    if kind == "rates":          # rates pass 72 in + 72 out, stress update ~370 B
        return n * (72.0 + 72.0 + 370.0 + lst)
    raise ValueError(kind)


def cpu_baseline(case, steps):
    sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
    from oracle_case import oracle_from_case

    threads = min(16, os.cpu_count() or 1)
    d = oracle_from_case(case, threads=threads)
    d.solve_init(case.dt)
    t0 = time.perf_counter()
    for _ in range(steps):
        d.step()
    wall = time.perf_counter() - t0
    return {"value": case.n * steps / wall, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"oracle (C+OpenMP restatement of SolveDiffUpdateLeapFrog, Nishimura order) "
                      f"on the same {case.n}-particle case, first {steps} steps incl. the initial "
                      f"neighbour search, {wall:.1f} s wall"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    from weldform_amd import capi, cases

    case = cases.config2(steps=args.steps + args.warmup)
    log(f"bench: {case.name} n={case.n} dt={case.dt:.4e} warmup={args.warmup} steps={args.steps}")
    s = cases.make_solver(case, precision=capi.FP64)
    s.step(args.warmup)
    s.sync()
    s.set_timing(True)
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    times = {k: s.kernel_time(k) for k in ("accel", "density", "rates", "neighbours", "step")}
    pairs, maxnb = s.neighbour_stats()
    sc = s.scalars()
    nbar = pairs / case.n
    ps = case.n * args.steps / wall
    log(f"bench: wall {wall:.3f} s  {1e3 * wall / args.steps:.3f} ms/step  {ps:.4e} particle-steps/s")
    log("bench: kernel avg ms " + ", ".join(f"{k}={v[0]:.4f} ({v[1]})" for k, v in times.items()))
    dom = max(("accel", "density", "rates"), key=lambda k: times[k][0])
    dom_ms = times[dom][0]
    ach = algorithmic_bytes(dom, case.n, nbar) / (dom_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": ps,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{case.name}: cylinder upsetting R=30mm L=60mm dx=0.5537mm, JC, v_top=-1m/s",
            "particles": case.n,
            "precision": "fp64",
            "pairs_unordered": pairs // 2,
            "mean_neighbours": nbar,
            "max_neighbours": maxnb,
            "pair_evaluations_per_s": (pairs // 2) * args.steps / wall,
            "kernel_ms": {k: v[0] for k, v in times.items()},
            "nb_searches": sc["nb_searches"],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu:
        ccase = cases.config2(steps=args.cpu_steps)
        log(f"bench: cpu baseline, {args.cpu_steps} steps")
        out["cpu_baseline"] = cpu_baseline(ccase, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
