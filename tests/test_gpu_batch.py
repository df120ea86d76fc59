"""Batched ensembles (several specimens advanced as one particle set, each
with its own material, no pair across specimens) against the CPU oracle and
against the same specimens run one by one.

The reference's neighbour-rebuild schedule is global: once any particle's
plastic strain passes 1e-6, every particle's list is rebuilt every
ts_nb_inc steps (Source/SolverLeapfrog.cpp:160-201).  In a batch the first
specimen to yield sets the schedule for all, so a specimen in a batch and
the same specimen alone can use neighbour lists of different ages; the
oracle, run on the whole batch, follows the batch's schedule.

Bars: FP64 batch vs oracle displacement and plastic strain within 1e-8
normwise after the run, per specimen; batch vs the specimen alone (FP64 and
mixed) within 1e-4 (the list ages differ); mixed batch vs the FP64 oracle
within 1e-4; the autotuner's candidates give bitwise-equal results for every
staging capacity."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import autotune, capi, cases

pytestmark = pytest.mark.gpu

STEPS = 120


@pytest.fixture(scope="module")
def batch():
    c = cases.config4(n_spec=6, n_range=(1500, 5000), steps=STEPS, v_top=-20.0, seed=5)
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    for _ in range(STEPS):
        d.step()
    return c, d


def test_batch_matches_oracle_and_lone_specimens(batch):
    c, d = batch
    print(f"batch: n={c.n} specimens={len(c.materials)} steps={STEPS}", flush=True)
    assert np.abs(d.f("PLSTRAIN")).max() > 0
    res = {}
    for prec in (capi.FP64, capi.MIXED):
        g = cases.make_solver(c, precision=prec)
        g.step(STEPS)
        g.sync()
        res[prec] = (g.get("disp"), g.get("pl_strain"))
        g.close()
    for i in range(len(c.materials)):
        sel = c.material == i
        s = cases.specimen(c, i)
        e_o = normwise_err(res[capi.FP64][0][sel], d.f("DISP")[sel])
        e_p = normwise_err(res[capi.FP64][1][sel], d.f("PLSTRAIN")[sel])
        e_mo = normwise_err(res[capi.MIXED][0][sel], d.f("DISP")[sel])
        lone = {}
        for prec in (capi.FP64, capi.MIXED):
            g = cases.make_solver(s, precision=prec)
            g.step(STEPS)
            g.sync()
            lone[prec] = g.get("disp")
            g.close()
        e_l = normwise_err(res[capi.FP64][0][sel], lone[capi.FP64])
        e_ml = normwise_err(res[capi.MIXED][0][sel], lone[capi.MIXED])
        print(f"specimen {i}: n={s.n} vs oracle {e_o:.2e} pl {e_p:.2e} mixed {e_mo:.2e}; "
              f"vs lone fp64 {e_l:.2e} mixed {e_ml:.2e}", flush=True)
        assert e_o < 1e-8 and e_p < 1e-8 and e_l < 1e-4
        assert e_mo < 1e-4 and e_ml < 1e-4


def test_autotune_shapes_agree(batch):
    c, _ = batch
    best, table = autotune.autotune(c, precision=capi.MIXED, steps=5, warmup=2, lanes=(2,), stages=(-1, 512, 0))
    print("autotune", best, table, flush=True)
    assert best in [{"lanes": r["lanes"], "stage_max": r["stage_max"]} for r in table]
    out = []
    for st in (-1, 512, 0):
        g = cases.make_solver(c, precision=capi.MIXED, lanes=2, stage_max=st)
        g.step(10)
        g.sync()
        out.append(g.get("disp"))
        g.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
