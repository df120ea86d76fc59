"""Launch-shape autotuner for a case on the current GPU.

The pair passes have two shape knobs that leave the physics alone: the
threads per particle (`lanes`: 1, 2 or 4; partial sums are combined in lane
order, so a choice fixes the rounding of the sums but no choice is more
exact than another) and, in mixed precision, the LDS staging capacity per
128-particle tile (`stage_max`: -1 gathers from HBM, 0 the largest capacity,
otherwise the largest staged set; results are bitwise the same for every
value).  `autotune`
times a short run of the case for each candidate shape, after a warm-up that
covers the first neighbour search, and returns the fastest.
"""
import time

from . import capi, cases

LANES = (1, 2, 4)
STAGES = (-1, 512, 0)


def autotune(case, precision=capi.MIXED, steps=20, warmup=5, lanes=LANES, stages=STAGES, log=None):
    """Returns (best, table): best = dict(lanes=..., stage_max=...), table =
    a list of dict(lanes, stage_max, ms_per_step) for every candidate."""
    table = []
    for ln in lanes:
        for st in (stages if precision == capi.MIXED else (-1,)):
            s = cases.make_solver(case, precision=precision, lanes=ln, stage_max=st)
            s.step(warmup)
            s.sync()
            t0 = time.perf_counter()
            s.step(steps)
            s.sync()
            ms = 1e3 * (time.perf_counter() - t0) / steps
            s.close()
            row = {"lanes": ln, "stage_max": st, "ms_per_step": ms}
            table.append(row)
            if log:
                log(f"autotune {case.name}: {row}")
    best = min(table, key=lambda r: r["ms_per_step"])
    return {"lanes": best["lanes"], "stage_max": best["stage_max"]}, table
