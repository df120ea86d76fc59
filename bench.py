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
# pair-pass kernels of the mixed-precision solver (two lanes per particle)
DOM_KERNEL = {"accel": "k_accel_mx<2>", "density": "k_dens_mx<2>", "rates": "k_rates_mx<2>"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kind, n, nbar, mixed, idx_bytes):
    """Algorithmic HBM bytes of one launch of a pair kernel: each particle's
    pair-pass inputs read once and outputs written once (8 B per gathered
    value in FP64, 4 B in mixed mode; outputs are FP64), its neighbour row
    read once (idx_bytes per entry: 4 for the global-index rows, 2 for the
    LDS-position rows of the staged mixed passes) plus the 4 B count, and
    the per-particle state the kernel fuses in.
      accel:   x(3) h v(3) m rho sigma(6) = 15 in; a(3) out
      density: x(3) h v(3) m m/rho = 9 in; drho out; fused drift: rho, disp
               in, rho, m/rho, x, disp out (+ the FP32 x shadow in mixed mode)
      rates:   x(3) h v(3) m rho = 9 in; strain rate (6) + rotation (3) out"""
    w = 4.0 if mixed else 8.0
    lst = idx_bytes * nbar + 4.0
    if kind == "accel":
        return n * (15 * w + 24.0 + lst)
    if kind == "density":
        return n * (9 * w + 8.0 + (8.0 + 24.0) + (8.0 + 8.0 + 24.0 + 24.0) + (16.0 if mixed else 0.0) + lst)
    if kind == "rates":
        return n * (9 * w + 72.0 + lst)
    raise ValueError(kind)


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, calibrated on a 1 GiB copy), or
    None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "traffic_cyl_upset_976k.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes" and parts[1] == kernel:
                return float(parts[2])
    return None


def profiled_valu(kernel):
    """(VALU issue share, VALU wave-instructions per launch) of `kernel` from
    the committed rocprofv3 counter summary, or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "pair_pass_counters_cyl_upset_976k.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == "valu_issue" and parts[1] == kernel:
                return float(parts[2]), float(parts[3])
    return None


def timed_run(case, precision, steps, warmup):
    """Graph-replay timed region for the throughput, then the same steps again
    launched one by one with HIP events around each pair kernel (per-kernel
    durations for the roofline)."""
    from weldform_amd import cases

    w = cases.make_solver(case, precision=precision)
    w.step(warmup)
    w.sync()
    w.close()
    ti = time.perf_counter()
    s = cases.make_solver(case, precision=precision)
    s.prepare()
    s.sync()
    init_s = time.perf_counter() - ti
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    e = cases.make_solver(case, precision=precision)
    e.sync()
    e.set_timing(True)
    t1 = time.perf_counter()
    e.step(steps)
    e.sync()
    wall_eager = time.perf_counter() - t1
    times = {k: e.kernel_time(k) for k in ("accel", "density", "rates", "neighbours", "step")}
    e.close()
    return s, wall, wall_eager, times, init_s

def yield_run(steps, chunk=50):
    """The headline billet (config 2 geometry and material, 976,320
    particles) upset at -20 m/s so that it yields: once a particle's plastic
    strain passes 1e-6 the reference rebuilds every neighbour list every
    ts_nb_inc = 5 steps (Source/SolverLeapfrog.cpp:160-201,347-356).  Mixed
    precision, graph replay, timed in chunks: the ms/step before and after
    the first yield show what the rebuilds cost."""
    import numpy as np

    from weldform_amd import capi, cases

    case = cases.config2(steps=steps)
    case.bcs = [dict(b) for b in case.bcs]
    for b in case.bcs:
        if b["zone"] == case.top_zone:
            b["value"] = (0.0, 0.0, -20.0)
    log(f"bench: yielding {case.name} n={case.n} v_top=-20 m/s steps={steps}")
    s = cases.make_solver(case, precision=capi.MIXED)
    s.prepare()
    s.sync()
    walls, searches, yielded_at, staging = [], [], None, []
    done = 0
    while done < steps:
        k = min(chunk, steps - done)
        t0 = time.perf_counter()
        s.step(k)
        s.sync()
        walls.append((k, time.perf_counter() - t0))
        sc = s.scalars()
        searches.append(sc["nb_searches"])
        if yielded_at is None and sc["max_pl_strain"] > 1e-6:
            yielded_at = done + k
        done += k
        if done == k or done == steps:
            # tiles whose staged set still fits LDS after this chunk's last build
            tiles, staged, values = s.stage_stats()
            staging.append({"after_step": done, "staged_tile_fraction": staged / tiles,
                            "mean_staged_set": values / max(staged, 1)})
    per = [searches[0]] + [b - a for a, b in zip(searches[:-1], searches[1:])]  # searches in each chunk
    pre = [w / k for (k, w), m in list(zip(walls, per))[1:] if m == 0]
    post = [w / k for (k, w), m in zip(walls, per) if m >= k // 5]
    s.close()
    res = {"workload": f"{case.name} at v_top=-20 m/s: 976,320 particles, mixed, {steps} Leapfrog steps, "
                       f"timed in chunks of {chunk}",
           "steps": steps, "nb_searches": int(searches[-1]), "yielded_by_step": yielded_at,
           "ms_per_step_chunks": [1e3 * w / k for k, w in walls], "searches_per_chunk": [int(m) for m in per],
           "ms_per_step_before_yield": 1e3 * float(np.median(pre)) if pre else None,
           "ms_per_step_yielding": 1e3 * float(np.median(post)) if post else None,
           "lds_staging": staging}
    log("bench: yielding " + json.dumps(res))
    return res


def contact_run(steps, warmup=10):
    """Thermo-mechanical contact compression (weldform_amd.cases.config3,
    384,256 particles, two rigid planes of 2,048 triangles, conduction and
    plastic heating), mixed precision, graph replay.  After the timed run,
    properties the physics fixes: the heat balance sum m c_p (T_N + dt
    dT/dt_N - T_0) = 0.9 x plastic work + contact heat (ThermalCalcs
    integrates T with the previous step's dT/dt), the signs of the planes'
    contact forces, a finite state."""
    import numpy as np

    from weldform_amd import capi, cases

    case = cases.config3(steps=steps)
    log(f"bench: {case.name} n={case.n} steps={steps}")
    w = cases.make_solver(case, precision=capi.MIXED)
    w.step(warmup)
    w.sync()
    w.close()
    s = cases.make_solver(case, precision=capi.MIXED)
    s.prepare()
    T0 = s.get("T")
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    wall = time.perf_counter() - t0
    sc = s.scalars()
    F = np.stack([s.contact_force_history(0), s.contact_force_history(1)], axis=1)
    T, dTdt = s.get("T"), s.get("dTdt")
    cp = case.materials[0]["cp_T"]
    gain = float(np.sum(case.mass * cp * (T + case.dt * dTdt - T0)))
    heat = 0.9 * sc["plastic_work"] + sc["accum_cont_heat_cond"] + sc["contact_friction_work"]
    finite = bool(np.isfinite(T).all() and np.isfinite(s.get("disp")).all())
    res = {"workload": f"{case.name}: R=15mm L=30mm dx=0.3757mm JC steel, thermal + plastic heat, "
                       f"top plane -5 m/s, bottom plane fixed, mu=0.2, {steps} Leapfrog steps from t=0, mixed",
           "particles": case.n, "steps": steps, "ms_per_step": 1e3 * wall / steps,
           "particle_steps_per_s": case.n * steps / wall, "contact_pairs": sc["contact_pairs"],
           "top_plane_force_z_mean_last_200": float(F[-200:, 0, 2].mean()),
           "bottom_plane_force_z_mean_last_200": float(F[-200:, 1, 2].mean()), "max_T": sc["max_T"],
           "plastic_work": sc["plastic_work"], "max_pl_strain": sc["max_pl_strain"],
           "nb_searches": sc["nb_searches"],
           "checks": {"heat_gain_J": gain, "heat_sources_J": heat,
                      "heat_balance_rel_err": abs(gain - heat) / heat if heat > 0 else None,
                      "plane_force_signs_ok": bool(F[-200:, 0, 2].mean() < 0 < F[-200:, 1, 2].mean()),
                      "finite_state": finite}}
    log("bench: config3 " + json.dumps(res))
    s.close()
    return res


def adjoint_run(case, kind, index, window, pre=100, k=20, fd=False, rel_step=1e-6, desc=""):
    """Adjoint calibration window: after `pre` steps, a `window`-step FP64
    Leapfrog window with a snapshot every k steps; loss = sum_n (F_z,n -
    target_n)^2 over a rigid plane's contact force (kind ADJ_PLANE) or a
    velocity-BC zone's reaction (ADJ_ZONE), target = 0.98 x the window's own
    history; one gradient = wf_adj_forward + wf_adj_backward
    (dL/d(A, B, n, C, m, friction)).  fd: central differences of the five JC
    parameters with a relative step rel_step, compared as theta dL/dtheta."""
    import numpy as np

    from weldform_amd import capi, cases

    log(f"bench: {case.name} n={case.n} window={window} checkpoint_every={k}")
    s = cases.make_solver(case, precision=capi.FP64)
    s.step(pre)
    s.sync()
    _, hist = s.adj_forward(np.zeros(window), checkpoint_every=k, kind=kind, index=index, component=2)
    s.adj_restore()
    target = 0.98 * hist
    t0 = time.perf_counter()
    loss, _ = s.adj_forward(target, checkpoint_every=k, kind=kind, index=index, component=2)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = s.adj_backward()
    t_bwd = time.perf_counter() - t0
    s.adj_restore()
    s.sync()
    t0 = time.perf_counter()
    s.step(window)
    s.sync()
    t_plain = time.perf_counter() - t0
    res = {"workload": f"{case.name}: {desc}; window of {window} FP64 Leapfrog steps after {pre}, snapshot every {k}",
           "particles": case.n, "window_steps": window, "checkpoint_every": k,
           "gradients_per_s": 1.0 / (t_fwd + t_bwd),
           "forward_window_s": t_fwd, "backward_s": t_bwd, "plain_forward_s": t_plain,
           "backward_over_forward": t_bwd / t_plain,
           "gradient_over_forward": (t_fwd + t_bwd) / t_plain,
           "loss": loss, "grad": dict(zip(capi.ADJ_PARAMS, [float(x) for x in g]))}
    if fd:
        mat0 = dict(case.materials[0])
        rel = {}
        for p, key in enumerate(("jc_A", "jc_B", "jc_n", "jc_C", "jc_m")):
            if mat0[key] == 0:
                continue
            h = rel_step * abs(mat0[key])
            L = []
            for sgn in (1, -1):
                s.adj_restore()
                s.set_material(0, **dict(mat0, **{key: mat0[key] + sgn * h}))
                L.append(s.adj_forward(target, checkpoint_every=k, kind=kind, index=index, component=2)[0])
            s.set_material(0, **mat0)
            fdv = (L[0] - L[1]) / (2 * h)
            rel[key] = {"theta_fd": fdv * mat0[key], "theta_adjoint": float(g[p]) * mat0[key],
                        "rel_err": abs(g[p] - fdv) / abs(fdv) if fdv else None}
        res["fd_check"] = {"rel_step": rel_step, "params": rel,
                           "max_rel_err": max(v["rel_err"] for v in rel.values() if v["rel_err"] is not None)}
    log("bench: adjoint " + json.dumps(res))
    s.close()
    return res


def adjoint_legs(window):
    """Timing at 200k (config 5 with contact, and the same billet under
    velocity BCs), plus finite-difference checks where the gradient is well
    posed: the 200k zone window with linear hardening (JC n = 1) and the
    4.5k aluminium billet over a 100-step window."""
    from weldform_amd import capi, cases

    out = {}
    c5 = cases.config5(steps=window)
    out["config5_adjoint"] = adjoint_run(c5, capi.ADJ_PLANE, 0, window, desc=(
        "R=15mm L=30mm dx=0.4675mm JC steel, isothermal, planes -5 m/s / fixed, mu=0.2, loss on the top "
        "plane's contact force"))
    cz = cases.config5_zone(steps=window)
    out["config5_adjoint_zone"] = adjoint_run(cz, capi.ADJ_ZONE, cz.top_zone, window, desc=(
        "the same billet, top row -5 m/s, bottom row fixed, loss on the top zone's reaction"))
    cl = cases.config5_zone(steps=window)
    cl.name = "cyl_upset_200k_jc_n1_adjoint"
    cl.materials = {0: dict(cl.materials[0], jc_n=1.0)}
    out["config5_adjoint_zone_n1_fd"] = adjoint_run(cl, capi.ADJ_ZONE, cl.top_zone, window, fd=True, desc=(
        "config5_zone with linear hardening (JC n = 1), finite-difference check over the whole window"))
    ca = cases.cylinder_case("cyl_upset_4k_jc_adjoint", 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"],
                             cases.AL["nu"], cases.JC_AL, steps=200, jitter=1e-3, seed=7, v_top=-20.0, T0=400.0)
    out["adjoint_4k_fd"] = adjoint_run(ca, capi.ADJ_ZONE, ca.top_zone, 100, fd=True, desc=(
        "4,524-particle aluminium billet, JC (n = 0.32), T0 = 400 K, top row -20 m/s, loss on the top zone's "
        "reaction"))
    return out


def batch_run(steps, n_sample=16):
    """Batched ensemble (weldform_amd.cases.config4: 256 cylinder specimens,
    N_i log-uniform on [2k, 40k], JC parameters +-20 %, each between its own
    rigid patches with its own Coulomb mu in [0.05, 0.3], 2.7M particles) in
    one context: shape chosen by the autotuner, then `steps` mixed-precision
    Leapfrog steps replayed from one captured graph per step.  One step
    advances every specimen, so each specimen's step latency is the batch
    step; reported per specimen are its share of the batch step (ms x
    N_i / N), its step time alone (a size-stratified sample of the specimens
    run alone, interpolated in N_i over all of them) and its top-patch force
    (per-specimen contact sums kept on the device)."""
    import numpy as np

    from weldform_amd import autotune, capi, cases

    case = cases.config4(steps=steps)
    nspec = len(case.materials)
    log(f"bench: {case.name} n={case.n} steps={steps}")
    best, table = autotune.autotune(case, precision=capi.MIXED, steps=10, warmup=3, lanes=(1, 2, 4),
                                    stages=(-1, 0), log=log)
    w = cases.make_solver(case, precision=capi.MIXED, **best)
    w.step(5)
    w.sync()
    w.close()
    s = cases.make_solver(case, precision=capi.MIXED, **best)
    s.prepare()
    s.sync()
    # timed in chunks of 50 steps: the chunks without a search (before the
    # specimens yield) against those with a rebuild every 5 steps
    walls, searches, done = [], [], 0
    t0 = time.perf_counter()
    while done < steps:
        k = min(50, steps - done)
        t1 = time.perf_counter()
        s.step(k)
        s.sync()
        walls.append((k, time.perf_counter() - t1))
        searches.append(s.scalars()["nb_searches"])
        done += k
    wall = time.perf_counter() - t0
    per = [searches[0]] + [b - a for a, b in zip(searches[:-1], searches[1:])]
    pre = [w / k for (k, w), m in list(zip(walls, per))[1:] if m == 0]
    post = [w / k for (k, w), m in zip(walls, per) if m >= k // 5]
    top = np.array([s.group_history(1, i, 0)[:, 2] for i in range(nspec)])  # (specimen, step)
    yielded = float((s.get("pl_strain") > 0).mean())
    # 20 more (yielding) steps launched one by one with HIP events around
    # each part of the step: where a yielding batch step goes
    nb0 = s.scalars()["nb_searches"]
    s.set_timing(True)
    s.step(20)
    s.sync()
    kt = {k: s.kernel_time(k)[0] for k in ("accel", "density", "rates", "neighbours", "surface", "step")}
    kt["searches"] = int(s.scalars()["nb_searches"] - nb0)
    s.set_timing(False)
    s.close()
    spec_n = case.meta["spec_n"]
    order = np.argsort(spec_n)
    pick = order[np.linspace(0, len(order) - 1, n_sample).round().astype(int)]
    lone = []
    for i in pick:
        sc = cases.specimen(case, int(i))
        g = cases.make_solver(sc, precision=capi.MIXED, **best)
        g.step(5)
        g.sync()
        t1 = time.perf_counter()
        g.step(50)
        g.sync()
        lone.append((int(sc.n), 1e3 * (time.perf_counter() - t1) / 50))
        g.close()
    ln = np.array([x[0] for x in lone], dtype=float)
    lt = np.array([x[1] for x in lone])
    alone = np.interp(spec_n, ln, lt)  # every specimen alone, step time interpolated in its size
    serial = float(alone.sum())
    ms = 1e3 * wall / steps
    share = ms * spec_n / spec_n.sum()
    tail = top[:, steps // 2:].mean(axis=1)
    res = {"workload": f"{case.name}: {nspec} cylinders, N_i log-uniform [2k, 40k], R/L uniform 0.2-0.5, "
                       f"dx=1mm, JC +-20 %, own rigid patches (top -5 m/s) and mu uniform 0.05-0.3 each, "
                       f"{steps} mixed Leapfrog steps, one context",
           "particles": case.n, "specimens": nspec, "steps": steps, "ms_per_step": ms,
           "particle_steps_per_s": case.n * steps / wall, "yielded_fraction": yielded,
           "ms_per_step_chunks": [1e3 * w / k for k, w in walls], "searches_per_chunk": [int(m) for m in per],
           "ms_per_step_before_yield": 1e3 * float(np.median(pre)) if pre else None,
           "ms_per_step_yielding": 1e3 * float(np.median(post)) if post else None,
           "kernel_ms_yielding_eager_20_steps": kt,
           "autotune": {"best": best, "table": table},
           "per_specimen_share_of_batch_step_ms": {"p50": float(np.percentile(share, 50)),
                                                   "p99": float(np.percentile(share, 99))},
           "per_specimen_step_ms_alone": {"p50": float(np.percentile(alone, 50)),
                                          "p99": float(np.percentile(alone, 99)), "sample": lone},
           "per_specimen_top_force_N_second_half_mean": {"min": float(tail.min()),
                                                         "p50": float(np.percentile(tail, 50)),
                                                         "max": float(tail.max())},
           "serial_ms_per_step_estimate": serial, "batched_speedup_vs_serial": serial / ms}
    log("bench: batch " + json.dumps(res))
    return res


def output_run(s, n, reps=3):
    """The reference's output files of the 976k state (Domain::WriteXDMF and
    Domain::WriteCSV): the tables packed on the device and copied to the
    host in one transfer each (best of `reps`), then written to a temporary
    directory (XDMF/HDF5 when the HDF5 library is present, CSV with 16
    formatting threads)."""
    import shutil
    import tempfile

    from weldform_amd import capi, output

    s.sync()
    res = {"particles": n, "xdmf_table_bytes": (capi.XDMF_NF + capi.XDMF_NI) * 4 * n,
           "csv_table_bytes": capi.CSV_NCOL * 8 * n}
    best = {}
    for name, fn in (("pack_xdmf", s.pack_xdmf), ("pack_csv", s.pack_csv)):
        for _ in range(reps):
            t0 = time.perf_counter()
            tab = fn()
            dt = time.perf_counter() - t0
            best[name] = min(best.get(name, dt), dt)
        if name == "pack_xdmf":
            f, iv = tab
        else:
            rows = tab
    res["pack_xdmf_ms"] = 1e3 * best["pack_xdmf"]
    res["pack_csv_ms"] = 1e3 * best["pack_csv"]
    res["pack_xdmf_GBps_host_inclusive"] = res["xdmf_table_bytes"] / best["pack_xdmf"] / 1e9
    d = tempfile.mkdtemp(prefix="wf_out_")
    try:
        if output.has_hdf5():
            t0 = time.perf_counter()
            output.write_xdmf_table(os.path.join(d, "bench_0001"), n, f, iv)
            res["write_xdmf_ms"] = 1e3 * (time.perf_counter() - t0)
        t0 = time.perf_counter()
        output.write_csv_rows(os.path.join(d, "bench_0001"), rows, threads=16)
        res["write_csv_ms"] = 1e3 * (time.perf_counter() - t0)
        res["file_bytes"] = {fn: os.path.getsize(os.path.join(d, fn)) for fn in sorted(os.listdir(d))}
    finally:
        shutil.rmtree(d, ignore_errors=True)
    log("bench: output " + json.dumps(res))
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps2, steps1=300, reps=3):
    """The oracle (C + OpenMP restatement of SolveDiffUpdateLeapFrog in the
    reference's Nproc = 1 pair order) on the host cores: config 2 (the
    headline workload) and config 1.  Step 1 (with the initial neighbour
    search) runs untimed; then `reps` runs of the step loop are timed with
    the steady clock and the median is reported, as the reference's step
    loop would be timed (set-up excluded)."""
    sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")]
    import statistics

    from oracle_case import oracle_from_case
    from weldform_amd import cases

    env = os.environ.get("OMP_NUM_THREADS")
    threads = int(env) if env and env.isdigit() else len(os.sched_getaffinity(0))
    out = {}
    for name, case, k in (("config2", cases.config2(steps=steps2), steps2),
                          ("config1", cases.config1(steps=steps1), steps1)):
        d = oracle_from_case(case, threads=threads)
        d.solve_init(case.dt)
        d.step()
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            for _ in range(k):
                d.step()
            walls.append(time.perf_counter() - t0)
        wall = statistics.median(walls)
        pairs = len(d.pairs())
        out[name] = {"particles": case.n, "steps": k, "median_wall_s": wall, "walls_s": walls,
                     "particle_steps_per_s": case.n * k / wall, "pair_evaluations_per_s": pairs * k / wall,
                     "ms_per_step": 1e3 * wall / k, "pairs_unordered": pairs}
        log(f"bench: cpu baseline {name} " + json.dumps(out[name]))
        del d
    c2 = out["config2"]
    return {"value": c2["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(),
            "sample": f"oracle (C+OpenMP restatement of SolveDiffUpdateLeapFrog, Nishimura order, "
                      f"gcc -O3) on the {c2['particles']}-particle headline case: step 1 (initial search) "
                      f"untimed, then the median of {reps} timed runs of {steps2} steps "
                      f"({c2['median_wall_s']:.2f} s); config 1 (4,524 particles) median of {reps} x "
                      f"{steps1} steps beside it",
            "configs": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="timed steps per run of the CPU baseline on the headline case (median of 3 runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-output", action="store_true", help="skip the output-file leg")
    ap.add_argument("--config3-steps", type=int, default=2000,
                    help="steps of the thermo-mechanical contact case reported beside the headline (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=500,
                    help="steps of the batched-ensemble leg (0: skip)")
    ap.add_argument("--adjoint-steps", type=int, default=400,
                    help="window of the adjoint calibration leg (0: skip)")
    ap.add_argument("--yield-steps", type=int, default=500,
                    help="steps of the yielding headline-billet leg (-20 m/s; 0: skip)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 1 GiB (for scaling FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        nbytes = 1 << 30
        a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is described in DESIGN.md but not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from weldform_amd import capi, cases

    case = cases.config2(steps=args.steps)
    log(f"bench: {case.name} n={case.n} dt={case.dt:.4e} warmup={args.warmup} steps={args.steps}")
    runs = {}
    for prec, name in ((capi.FP64, "fp64"), (capi.MIXED, "mixed")):
        s, wall, wall_eager, times, init_s = timed_run(case, prec, args.steps, args.warmup)
        runs[name] = dict(solver=s, wall=wall, wall_eager=wall_eager, times=times, init_s=init_s,
                          R=s.reaction_history(case.top_zone)[:, 2], disp=s.get("disp"),
                          ps=case.n * args.steps / wall)
        log(f"bench: {name}: {1e3 * wall / args.steps:.3f} ms/step (graph), "
            f"{1e3 * wall_eager / args.steps:.3f} ms/step (eager, timed)  {runs[name]['ps']:.4e} particle-steps/s  "
            "kernel avg ms " + ", ".join(f"{k}={v[0]:.4f}" for k, v in times.items()))
    f64, mx = runs["fp64"], runs["mixed"]
    pairs, maxnb = mx["solver"].neighbour_stats()
    sc = mx["solver"].scalars()
    nbar = pairs / case.n

    def nerr(a, b):
        return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))

    err = {"mixed_vs_fp64_gpu_reaction_history": nerr(mx["R"], f64["R"]),
           "mixed_vs_fp64_gpu_displacement": nerr(mx["disp"], f64["disp"]),
           "mixed_vs_fp64_gpu_final_reaction": abs(mx["R"][-1] - f64["R"][-1]) / abs(f64["R"][-1])}
    gold = os.path.join(ROOT, "tests", "golden", "config2_oracle.npz")
    if os.path.exists(gold) and args.steps == 1000:
        g = np.load(gold)
        if int(g["n"]) == case.n and int(g["steps"]) == args.steps:
            for name, r in runs.items():
                err[f"{name}_vs_oracle_reaction_history"] = nerr(r["R"], g["R3"][:, 2])
                err[f"{name}_vs_oracle_displacement_subsample"] = nerr(r["disp"][g["idx"]], g["disp"])
    log("bench: errors " + json.dumps(err))
    times = mx["times"]
    dom = max(("accel", "density", "rates"), key=lambda k: times[k][0])
    dom_ms = times[dom][0]
    ach = algorithmic_bytes(dom, case.n, nbar, True, 2.0) / (dom_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": mx["ps"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * mx["wall"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair passes + f64 integration/return map/reductions",
        "data": "synthetic",
        "config": {
            "workload": f"{case.name}: cylinder upsetting R=30mm L=60mm dx=0.5537mm (976,320 particles), "
                        f"JC, v_top=-1m/s, {args.steps} Leapfrog steps from t=0",
            "particles": case.n,
            "precision": "mixed",
            "fp64_particle_steps_per_s": f64["ps"],
            "fp64_ms_per_step": 1e3 * f64["wall"] / args.steps,
            "mixed_speedup_vs_fp64": mx["ps"] / f64["ps"],
            "errors": err,
            "pairs_unordered": pairs // 2,
            "mean_neighbours": nbar,
            "max_neighbours": maxnb,
            "pair_evaluations_per_s": (pairs // 2) * args.steps / mx["wall"],
            "kernel_ms_mixed": {k: v[0] for k, v in times.items()},
            "kernel_ms_fp64": {k: v[0] for k, v in f64["times"].items()},
            "eager_ms_per_step_mixed": 1e3 * mx["wall_eager"] / args.steps,
            "eager_ms_per_step_fp64": 1e3 * f64["wall_eager"] / args.steps,
            "timing_note": "value: steps replayed from one captured HIP graph per step; kernel_ms: "
                           "the same steps launched one by one with HIP events on the solver's stream",
            "nb_searches": sc["nb_searches"],
            "init_seconds": {"mixed": mx["init_s"], "fp64": f64["init_s"]},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom + " (mixed)",
            "achieved": ach,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS,
            "traffic": profiled_traffic(DOM_KERNEL[dom]),
            "algorithmic_bytes": algorithmic_bytes(dom, case.n, nbar, True, 2.0),
            # the pair passes are VALU-bound: share of the SIMDs' issue cycles
            # their VALU instructions take, from the rocprofv3 counter summary
            # (profiles/pair_pass_counters_cyl_upset_976k.md)
            "valu_issue_frac": {k: (profiled_valu(v) or (None,))[0] for k, v in DOM_KERNEL.items()},
            "note": "achieved = algorithmic bytes of one launch / its mean duration (HIP events on the "
                    "solver stream); traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE "
                    "(profiles/traffic_cyl_upset_976k.md); valu_issue_frac = VALU issue cycles / SIMD cycles "
                    "at 2.4 GHz (profiles/pair_pass_counters_cyl_upset_976k.md)",
        },
    }
    if not args.no_output:
        out["config"]["output_976k"] = output_run(mx["solver"], case.n)
    if args.yield_steps > 0:
        out["config"]["yielding_976k"] = yield_run(args.yield_steps)
    if args.config3_steps > 0:
        out["config"]["config3_contact_thermal"] = contact_run(args.config3_steps)
    if args.batch_steps > 0:
        out["config"]["config4_batch"] = batch_run(args.batch_steps)
    if args.adjoint_steps > 0:
        out["config"].update(adjoint_legs(args.adjoint_steps))
    if not args.no_cpu:
        log(f"bench: cpu baseline, {args.cpu_steps} steps x 3")
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
