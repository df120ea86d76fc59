"""The oracle's contact heat (Domain::cont_heat_cond / cont_heat_fric,
Source/Contact_Wang.cpp:245-324, Source/Thermal.cpp:165-210) and contact
time-step criterion (auto_ts_cont, Source/Domain.cpp:222-229,
Source/Contact_Wang.cpp:226-233), pinned on the CPU by what the physics
fixes rather than by the formulas themselves:

  * energy: with plastic heating off and the particles' own conduction
    conservative (its pair terms cancel in sum m cp dT/dt), the heat the
    billet gains in a step, sum m cp (T' - T), equals the previous step's
    increments of the conducted and friction heat totals (ThermalCalcs
    integrates T with the dT/dt of the step before, Source/Thermal.cpp:
    721-748, and the step size is fixed here);
  * direction and area: heat flows from the hot punch into the billet, each
    contact particle through the area (m / rho)^(2/3) (Fraser 3-119), so
    q m / rho / (h_c (T_punch - T)) recovers it;
  * Coulomb: on sliding particles the contact shear stress is the dynamic
    coefficient (not the static one) times the normal pressure, and the
    friction heat is carried by the same area;
  * the contact criterion only ever shortens the step, and when it acts the
    step is Domain::min_force_ts = 2 m |v| / |f_n| of a contact particle
    (the billet moves up into the punch at 0.2 m/s: a particle at rest when
    it touches gives dt_fext = 0, the reference's step then drops to 0 and
    the next contact force is not finite)."""
import numpy as np

from oracle_case import oracle_from_case
from weldform_amd import cases

HC = 5.0e4
MU_S, MU_D = 0.3, 0.2


def _case(steps=120):
    c = cases.contact_small(steps=steps, thermal=True)
    c.pl_heating = False
    c.contact = (MU_S, MU_D, 0.0)
    c.contact_heat = (True, True, HC)
    nr = 2 * 22 ** 2
    c.rigid_T = np.concatenate([np.full(nr, 400.0), np.full(nr, 300.0)])
    return c


def test_contact_heat_energy_direction_and_coulomb():
    c = _case()
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    n = c.n
    cp = c.materials[0]["cp_T"]
    checked, dw_prev = 0, 0.0
    for s in range(c.steps):
        T0 = d.f("T")[:n].copy()
        w0 = d.param("contact_friction_work") + d.param("accum_cont_heat_cond")
        d.step()
        gain = float(np.sum(d.f("MASS")[:n] * cp * (d.f("T")[:n] - T0)))
        if dw_prev != 0.0:
            # the particles' conduction cancels to rounding of its (larger) flows
            assert abs(gain - dw_prev) <= 1e-6 * abs(dw_prev), (s, gain, dw_prev)
            checked += 1
        dw_prev = d.param("contact_friction_work") + d.param("accum_cont_heat_cond") - w0
    assert checked > 20
    qc, qf = d.f("QCONT")[:n], d.f("QFRIC")[:n]
    cs, fh = d.f("CSHEAR")[:n], d.f("FHFL")[:n]
    m, rho, T = d.f("MASS")[:n], d.f("RHO")[:n], d.f("T")[:n]
    area = (m / rho) ** (2.0 / 3.0)
    zmid = 0.5 * (c.x[:, 2].min() + c.x[:, 2].max())
    top = (qc != 0) & (d.f("X")[:n, 2] > zmid)
    assert top.sum() > 50                        # the punch face is in contact
    assert (qc[top] > 0).all()                   # heat flows into the billet
    a_rec = qc[top] * m[top] / rho[top] / (HC * (400.0 - T[top]))
    assert np.allclose(a_rec, area[top], rtol=2e-3)
    # Coulomb: contact shear = mu_dyn x normal pressure on sliding particles
    cf = d.f("CONTFORCE")[:n]
    slide = cs > 0
    assert slide.sum() > 20 and (qf[slide] >= 0).all()
    assert np.allclose(cs[slide] * area[slide], MU_D * np.abs(cf[slide, 2]), rtol=2e-3)
    assert np.allclose(fh[slide] * area[slide], qf[slide] * m[slide] / rho[slide], rtol=2e-3)
    assert d.param("contact_friction_work") > 0 and d.param("accum_cont_heat_cond") > 0


def auto_ts_cont_case(steps=80):
    c = cases.contact_small(steps=steps)
    c.v = np.tile([0.0, 0.0, 0.2], (c.n, 1))
    return c


def test_contact_time_step_criterion():
    c = auto_ts_cont_case()
    c.auto_ts = 1
    a = oracle_from_case(c, threads=8)
    a.solve_init(c.dt)
    c.auto_ts = 5
    b = oracle_from_case(c, threads=8)
    b.solve_init(c.dt)
    acted = 0
    for s in range(c.steps):
        a.step()
        b.step()
        da, db = a.param("deltat"), b.param("deltat")
        if db != da:
            acted += 1
            assert db == b.param("min_force_ts") and 0 < db < da, (s, da, db)
            # the least 2 m |v| / |f_n| over the particles the punch pushes
            assert (np.abs(b.f("CONTFORCE")[:c.n, 2]) > 0).any()
            break
        assert np.array_equal(a.f("X"), b.f("X"))
    assert acted == 1   # the punch's contact forces shorten the step within the run
