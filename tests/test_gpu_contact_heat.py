"""Contact heat and the contact time-step criterion on the GPU against the
CPU oracle (Source/Contact_Wang.cpp:226-233, 245-324; Source/Thermal.cpp:
165-210; Source/Domain.cpp:222-229).

  * Contact heat: the thermal contact case (a punch at 400 K on a 300 K
    billet, h_c 5e4 W/m^2K, friction heating, static 0.3 / dynamic 0.2)
    over the first 55 steps (contact starts at step ~45): the billet's
    heating T - 300, displacement and contact forces, the XDMF q_friction /
    c_shearabs columns, and the two heat totals.  Bars as the other contact
    runs: the penalty contact is chaotic, so max(1e-8, 10x the oracle's own
    sensitivity to a 1e-15 relative change of the initial positions) in
    FP64 (that sensitivity is 1.5e-9 - 4e-9 at step 55 and grows past 1e-5
    by step 70; measured: heating and the columns bit-equal, the rest
    <= 1e-12); the mixed path within 1e-4 (measured <= 1.1e-5).
  * auto_ts_cont: the velocity criterion with the contact criterion, the
    billet moving up into the punch, up to the step where the contact
    criterion first sets the step (45 steps): step size, time and
    displacement at the FP64 bars."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

HC = 5.0e4


def _heat_case(steps):
    c = cases.contact_small(steps=steps, thermal=True)
    c.pl_heating = False
    c.contact = (0.3, 0.2, 0.0)
    c.contact_heat = (True, True, HC)
    nr = 2 * 22 ** 2
    c.rigid_T = np.concatenate([np.full(nr, 400.0), np.full(nr, 300.0)])
    c.name = "cyl_upset_4k_contact_heat"
    return c


def _oracle(c, eps=0.0):
    d = oracle_from_case(c, threads=8)
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:c.n] = c.x * (1 + eps * rng.uniform(-1, 1, size=c.x.shape))
    d.solve_init(c.dt)
    for _ in range(c.steps):
        d.step()
    return d


def _col(f, n, off, w=1):
    return f[off * n:(off + w) * n]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_contact_heat_matches_oracle(precision):
    c = _heat_case(55)
    fp64 = precision == capi.FP64
    print(f"contact heat: n={c.n} steps={c.steps} precision={precision}", flush=True)
    d = _oracle(c)
    dp = _oracle(c, 1e-15)
    n = c.n

    def fields(get):
        return {"heat": get("T") - 300.0, "disp": get("disp"), "contforce": get("contforce"),
                "q_friction": get("fhfl"), "c_shearabs": get("cshear")}

    ref = fields(lambda k: {"T": d.f("T")[:n], "disp": d.f("DISP")[:n], "contforce": d.f("CONTFORCE")[:n],
                            "fhfl": d.f("FHFL")[:n], "cshear": d.f("CSHEAR")[:n]}[k])
    per = fields(lambda k: {"T": dp.f("T")[:n], "disp": dp.f("DISP")[:n], "contforce": dp.f("CONTFORCE")[:n],
                            "fhfl": dp.f("FHFL")[:n], "cshear": dp.f("CSHEAR")[:n]}[k])
    assert ref["q_friction"].max() > 0 and ref["heat"].max() > 0
    g = cases.make_solver(c, precision=precision)
    g.step(c.steps)
    g.sync()
    f, _ = g.pack_xdmf()
    got = {"heat": g.get("T") - 300.0, "disp": g.get("disp"), "contforce": g.get("contforce"),
           "q_friction": _col(f, n, capi.XDMF_Q_FRICTION).astype(np.float64),
           "c_shearabs": _col(f, n, capi.XDMF_C_SHEARABS).astype(np.float64)}
    ref32 = {k: (v.astype(np.float32).astype(np.float64) if k in ("q_friction", "c_shearabs") else v)
             for k, v in ref.items()}
    errs = {k: normwise_err(got[k], ref32[k]) for k in got}
    sens = {k: normwise_err(per[k], ref[k]) for k in ref}
    s = g.scalars()
    for k in ("contact_friction_work", "accum_cont_heat_cond"):
        errs[k] = abs(s[k] - d.param(k)) / abs(d.param(k))
        sens[k] = abs(dp.param(k) - d.param(k)) / abs(d.param(k))
    print("contact heat errors", errs, "oracle sensitivity", sens, flush=True)
    floor = 1e-8 if fp64 else 1e-4
    for k, e in errs.items():
        # float32 columns carry their own rounding
        fl = max(floor, 1e-6) if k in ("q_friction", "c_shearabs") else floor
        assert e < (max(fl, 10 * sens[k]) if fp64 else fl), k
    g.close()


def test_contact_time_step_criterion_matches_oracle():
    """Up to the step where the criterion first acts (the billet moves up
    into the punch at 0.2 m/s; afterwards the reference's criterion keeps
    shrinking the step, dt' = |v| dt^2 / delta, towards 0)."""
    c = cases.contact_small(steps=80)
    c.v = np.tile([0.0, 0.0, 0.2], (c.n, 1))
    c1 = cases.contact_small(steps=80)
    c1.v = c.v
    c.auto_ts, c1.auto_ts = 5, 1
    d, d1 = oracle_from_case(c, threads=8), oracle_from_case(c1, threads=8)
    d.solve_init(c.dt)
    d1.solve_init(c1.dt)
    k = 0
    while k < 80:
        d.step()
        d1.step()
        k += 1
        if d.param("deltat") != d1.param("deltat"):
            break
    assert 0 < d.param("deltat") < d1.param("deltat") and d.param("deltat") == d.param("min_force_ts")
    c.steps = k
    c.name = "cyl_upset_4k_contact_auto_ts_cont"
    print(f"auto_ts_cont: n={c.n} steps={c.steps}", flush=True)
    dp = _oracle(c, 1e-15)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(c.steps)
    g.sync()
    s = g.scalars()
    errs = {"dt": abs(s["dt"] - d.param("deltat")) / d.param("deltat"),
            "time": abs(s["time"] - d.param("Time")) / d.param("Time"),
            "disp": normwise_err(g.get("disp"), d.f("DISP")[:c.n])}
    sens = {"dt": abs(dp.param("deltat") - d.param("deltat")) / d.param("deltat"),
            "time": abs(dp.param("Time") - d.param("Time")) / d.param("Time"),
            "disp": normwise_err(dp.f("DISP")[:c.n], d.f("DISP")[:c.n])}
    print("auto_ts_cont errors", errs, "oracle sensitivity", sens, flush=True)
    for key, e in errs.items():
        assert e < max(1e-8, 10 * sens[key]), key
    g.close()
