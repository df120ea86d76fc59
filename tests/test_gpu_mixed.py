"""Mixed precision (FP32 pair arithmetic, FP64 integration, return map and
reductions): errors stated against the FP64 oracle.

Tolerances: single pair passes within 2e-5 normwise of the FP64 oracle on the
same inputs; reaction-force history within 1e-4 normwise after 1000 steps
(the target for the mixed step)."""
import os

import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

TOL_PASS_MIXED = 2e-5
TOL_RUN_MIXED = 1e-4


def _log(*a):
    print(*a, flush=True)


def test_mixed_single_passes():
    c = cases.config1()
    rng = np.random.default_rng(3)
    v = rng.normal(0.0, 1.0, size=(c.n, 3))
    s = rng.normal(0.0, 1e8, size=(c.n, 3, 3))
    sig = 0.5 * (s + np.transpose(s, (0, 2, 1)))
    g = cases.make_solver(c, precision=capi.MIXED)
    g.set("v", v)
    g.set("sigma", sig)
    g.rebuild_neighbours()
    d = oracle_from_case(c)
    d.f("V")[:] = v
    d.f("SIGMA")[:] = sig
    d.neighbour_search()
    d.start_acceleration()
    g.eval("accel")
    d.calc_accel()
    e_a = normwise_err(g.get("a"), d.f("A"))
    g.eval("density")
    d.calc_dens()
    e_d = normwise_err(g.get("ddensity"), d.f("DRHO"))
    g.eval("rates")
    d.calc_rates()
    e_s = normwise_err(g.get("strain_rate"), d.f("SR"))
    e_w = normwise_err(g.get("rot_rate"), d.f("RR"))
    _log("mixed pass errors accel", e_a, "density", e_d, "strain rate", e_s, "rotation", e_w)
    assert max(e_a, e_d, e_s, e_w) < TOL_PASS_MIXED


def test_mixed_config1_run():
    c = cases.config1()
    g = cases.make_solver(c, precision=capi.MIXED)
    g.step(c.steps)
    g.sync()
    d = oracle_from_case(c, threads=4)
    d.solve_init(c.dt)
    R = []
    for _ in range(c.steps):
        d.step()
        R.append(d.reaction(c.top_zone)[2])
    e_r = normwise_err(g.reaction_history(c.top_zone)[:, 2], np.array(R))
    e_u = normwise_err(g.get("disp"), d.f("DISP"))
    _log("mixed config1 reaction err", e_r, "disp err", e_u)
    assert e_r < TOL_RUN_MIXED and e_u < TOL_RUN_MIXED


def test_mixed_config2_vs_oracle_golden():
    gz = np.load(os.path.join(os.path.dirname(__file__), "golden", "config2_oracle.npz"))
    c = cases.config2(steps=int(gz["steps"]))
    g = cases.make_solver(c, precision=capi.MIXED)
    g.step(c.steps)
    g.sync()
    e_r = normwise_err(g.reaction_history(c.top_zone)[:, 2], gz["R3"][:, 2])
    e_f = abs(g.reaction_history(c.top_zone)[-1, 2] - gz["R3"][-1, 2]) / abs(gz["R3"][-1, 2])
    e_u = normwise_err(g.get("disp")[gz["idx"]], gz["disp"])
    _log("mixed config2 reaction history err", e_r, "final reaction err", e_f, "disp err", e_u)
    assert e_r < TOL_RUN_MIXED and e_f < TOL_RUN_MIXED and e_u < TOL_RUN_MIXED


def _mixed_run(case, steps, **kw):
    g = cases.make_solver(case, precision=capi.MIXED, **kw)
    g.step(steps)
    g.sync()
    return g


@pytest.mark.parametrize("stage_max", [0, 520])
def test_lds_staging_is_transparent(stage_max):
    """LDS-staged gathers read the same values as HBM gathers in the same
    order: the run must be bitwise identical with staging off (-1), on for
    every tile (0) and on for part of the tiles (520 = about half of them
    fall back to HBM).  v_top = -20 m/s yields early, so the staging lists
    are rebuilt with the neighbour lists every 5 steps."""
    c = cases.cylinder_case("cyl_upset_4k_jc_fast", 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"],
                            cases.AL["nu"], cases.JC_AL, steps=120, jitter=1e-3, seed=7, v_top=-20.0, T0=300.0)
    _log(f"staging transparency: stage_max={stage_max} n={c.n} steps={c.steps}")
    ref = _mixed_run(c, c.steps, stage_max=-1)
    got = _mixed_run(c, c.steps, stage_max=stage_max)
    assert ref.scalars()["nb_searches"] > 5
    assert got.scalars()["max_pl_strain"] > 0
    for f in ("x", "v", "a", "disp", "density", "sigma", "strain_rate", "rot_rate", "pl_strain"):
        assert np.array_equal(got.get(f), ref.get(f)), f
    assert np.array_equal(got.reaction_history(c.top_zone), ref.reaction_history(c.top_zone))
    assert got.scalars()["kin_energy_sum"] == ref.scalars()["kin_energy_sum"]
    # which tiles the last build staged: none with staging off; all with the
    # default bound (every set of this small billet fits); with 520 about
    # half, each staged set within the bound
    tiles, staged, values = got.stage_stats()
    assert tiles == -(-c.n // 128)
    assert ref.stage_stats()[1] == 0
    if stage_max == 0:
        assert staged == tiles
    else:
        assert 0 < staged < tiles
        assert values <= 520 * staged
