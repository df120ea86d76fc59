"""Thermo-mechanical contact compression as a whole on the GPU: rigid-plane
Wang contact with Coulomb friction, conduction and plastic-work heating in
one run (Source/SolverLeapfrog.cpp:215-330, Source/Contact_Wang.cpp:15-348,
Source/Thermal.cpp:53-219,721-748, Source/ThermalStuct.cpp:14-39).

  * Against the oracle (the steel billet of contact_small with the thermal
    solver and plastic heating on, mu = 0.2, the punch at -20 m/s so the
    billet touches the punch at step 13 and yields under it from step 19,
    200 steps): after 22 steps (contact, plastic strain, plastic heat and
    conduction all active) and after 200 steps the temperature rise, displacement, plastic strain and
    both plane forces within 1e-8 normwise in FP64 (1e-4 mixed), or within
    10x the oracle's own sensitivity to a 1e-15 relative change of its
    initial positions (1e-7 for the mixed path) where that is larger,
    because the penalty contact is chaotic once particles touch (see
    test_gpu_contact.py).
  * Energy over long runs, where the contact phase cannot be compared
    step by step: the heat the billet gains in a step, sum m c_p (T' - T),
    equals the previous step's increments of 0.9 x the plastic work plus
    the conducted and friction contact heat (ThermalCalcs integrates T with
    the previous step's dT/dt; the particles' own conduction cancels), held
    within 1e-5 relative at every step (the oracle holds it to 1.4e-7) over
    300 steps of the 4.5k billet with contact heat, and cumulatively over
    2000 steps of the 384k config-3 billet (mixed precision), together with
    the sign of the plane forces (the punch pushes the billet down, the
    bottom plane pushes it up) and a finite state."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

EARLY = 22
HC = 5.0e4
V_PUNCH = -20.0


def _oracle(c, eps=0.0):
    d = oracle_from_case(c, threads=8)
    n = c.n
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:n] = c.x * (1 + eps * rng.uniform(-1, 1, size=c.x.shape))
    d.solve_init(c.dt)
    F, early = [], None

    def snap():
        return {"heat": d.f("T")[:n] - 300.0, "disp": d.f("DISP")[:n].copy(), "pl": d.f("PLSTRAIN")[:n].copy()}

    for s in range(c.steps):
        d.step()
        F.append(np.stack([d.contact_force(0), d.contact_force(1)]))
        if s + 1 == EARLY:
            early = snap()
    out = snap()
    out["F"] = np.array(F)
    early["F"] = out["F"][:EARLY]
    return d, out, early


@pytest.fixture(scope="module")
def config3_small():
    c = cases.contact_small(steps=200, thermal=True, v_punch=V_PUNCH)
    c.name = "cyl_compress_4k_contact_thermal"
    d, ref, early = _oracle(c)
    assert c.pl_heating and c.thermal and c.contact[0] == 0.2
    # the run is not vacuous: plastic heat, conduction and both planes act
    assert d.f("QPL")[:c.n].max() > 0 and ref["heat"].max() > 0 and ref["pl"].max() > 1e-6
    assert np.abs(ref["F"][:, 0, 2]).max() > 0 and np.abs(ref["F"][:, 1, 2]).max() > 0
    sens = {}
    for eps in (1e-15, 1e-7):
        _, per, per_early = _oracle(c, eps)
        sens[eps] = (_errs(per, ref), _errs(per_early, early))
    return c, ref, early, sens


def _gpu_fields(g):
    return {"heat": g.get("T") - 300.0, "disp": g.get("disp"), "pl": g.get("pl_strain"),
            "F": np.stack([g.contact_force_history(0), g.contact_force_history(1)], axis=1)}


def _errs(got, ref):
    e = {k: normwise_err(got[k], ref[k]) for k in ("heat", "disp", "pl")}
    for m, name in ((0, "F_top"), (1, "F_bottom")):
        e[name] = normwise_err(got["F"][:, m, 2], ref["F"][:, m, 2])
    return e


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_config3_contact_conduction_plastic_heat_matches_oracle(config3_small, precision):
    c, ref, early, sens_all = config3_small
    fp64 = precision == capi.FP64
    print(f"config3 small: n={c.n} steps={c.steps} precision={precision}", flush=True)
    sens, sens_early = sens_all[1e-15 if fp64 else 1e-7]
    floor = 1e-8 if fp64 else 1e-4
    g = cases.make_solver(c, precision=precision)
    g.step(EARLY)
    g.sync()
    e_early = _errs(_gpu_fields(g), early)
    print("config3 early errors", e_early, "oracle sensitivity", sens_early, flush=True)
    for k, e in e_early.items():
        assert e < max(floor, 10 * sens_early[k]), k
    g.step(c.steps - EARLY)
    g.sync()
    errs = _errs(_gpu_fields(g), ref)
    print("config3 errors", errs, "oracle sensitivity", sens, flush=True)
    for k, e in errs.items():
        assert e < max(floor, 10 * sens[k]), k
    g.close()


def _heat_case(steps):
    c = cases.contact_small(steps=steps, thermal=True, v_punch=V_PUNCH)
    c.contact_heat = (True, True, HC)
    nr = 2 * 22 ** 2
    c.rigid_T = np.concatenate([np.full(nr, 400.0), np.full(nr, 300.0)])
    c.name = "cyl_compress_4k_contact_thermal_heat"
    return c


def _totals(g):
    s = g.scalars()
    return np.array([s["plastic_work"], s["accum_cont_heat_cond"], s["contact_friction_work"]])


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_config3_heat_balance_every_step(precision):
    c = _heat_case(300)
    print(f"config3 heat balance: n={c.n} steps={c.steps} precision={precision}", flush=True)
    cp = c.materials[0]["cp_T"]
    g = cases.make_solver(c, precision=precision)
    m = c.mass
    w = np.array([0.9, 1.0, 1.0])
    prev, checked, worst = None, 0, 0.0
    T = g.get("T")
    tot = _totals(g)
    for s in range(c.steps):
        g.step(1)
        g.sync()
        T1, tot1 = g.get("T"), _totals(g)
        gain = float(np.sum(m * cp * (T1 - T)))
        assert np.isfinite(T1).all()
        if prev is not None and prev > 0:
            r = abs(gain - prev) / prev
            worst = max(worst, r)
            assert r < 1e-5, (s, gain, prev)
            checked += 1
        prev = float(w @ (tot1 - tot))
        T, tot = T1, tot1
    print(f"config3 heat balance: {checked} steps checked, worst relative error {worst:.3e}, totals {tot}",
          flush=True)
    assert checked > 200
    assert tot[0] > 0 and tot[1] > 0 and tot[2] > 0 and T.max() > 300.0
    g.close()


def test_config3_384k_2000_steps_properties():
    """The config-3 billet (384,256 particles, two 2,048-triangle planes,
    conduction + plastic heating, mu 0.2, mixed precision) over 2000 steps:
    cumulative heat balance (T after the next step, T_N + dt dT/dt_N, against
    the totals through step N), plane-force signs, finite state."""
    c = cases.config3(steps=2000)
    cp = c.materials[0]["cp_T"]
    print(f"config3 384k: n={c.n} steps={c.steps} mixed", flush=True)
    g = cases.make_solver(c, precision=capi.MIXED)
    T0 = g.get("T")
    g.step(c.steps)
    g.sync()
    tot = _totals(g)
    T, dTdt = g.get("T"), g.get("dTdt")
    assert np.isfinite(T).all() and np.isfinite(g.get("disp")).all()
    gain = float(np.sum(c.mass * cp * (T + c.dt * dTdt - T0)))
    want = 0.9 * tot[0] + tot[1] + tot[2]
    F = np.stack([g.contact_force_history(0), g.contact_force_history(1)], axis=1)
    s = g.scalars()
    print(f"config3 384k: heat gain {gain:.6e} J, 0.9 plastic work + contact heat {want:.6e} J, "
          f"max T {s['max_T']:.4f}, max pl {s['max_pl_strain']:.4e}, F_top_z {F[-1, 0, 2]:.4e}, "
          f"F_bottom_z {F[-1, 1, 2]:.4e}, contact pairs {s['contact_pairs']}", flush=True)
    assert want > 0 and abs(gain - want) < 1e-5 * want
    assert s["max_pl_strain"] > 1e-4 and s["max_T"] > 300.0
    # the planes' contact forces on the billet (Domain::m_contact_force):
    # the driven punch pushes it down, the fixed bottom plane up
    assert F[-200:, 0, 2].mean() < 0 < F[-200:, 1, 2].mean()
    g.close()
