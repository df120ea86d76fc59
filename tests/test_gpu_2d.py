"""Plane strain (wf_set_dimension(2), Domain::Dimension 2) on the GPU
against the oracle.

Bars: single FP64 passes within 1e-10 normwise (the 3D bar); FP64 runs
within 1e-8 normwise, widened to 10x the oracle's own sensitivity to a
1e-15 relative perturbation of the positions for the fields that switch
with yield; mixed-precision runs within 1e-4 normwise (the mixed bar) or
10x the oracle's sensitivity to a 1e-7 perturbation.  In every 2D run the
z coordinate, the z velocity and the z acceleration stay exactly 0."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


def _log(*a):
    print(*a, flush=True)


def _oracle_run(c, scheme="leapfrog", eps=0.0, params=None):
    d = oracle_from_case(c, threads=8)
    for k, v in (params or {}).items():
        d.set_param(k, v)
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:, :2] = c.x[:, :2] * (1 + eps * rng.uniform(-1, 1, size=(c.n, 2)))
    d.solve_init(c.dt)
    R = []
    for _ in range(c.steps):
        d.step(scheme)
        R.append(d.reaction(c.top_zone).copy())
    return d, np.array(R)


def _fields(g, d, R_g, R_o):
    return {"reaction": normwise_err(R_g[:, 1], R_o[:, 1]),
            "disp": normwise_err(g.get("disp"), d.f("DISP")),
            "v": normwise_err(g.get("v"), d.f("V")),
            "sigma": normwise_err(g.get("sigma"), d.f("SIGMA")),
            "pl_strain": normwise_err(g.get("pl_strain"), d.f("PLSTRAIN"))}


def _sens(c, d, R, eps, scheme="leapfrog", params=None):
    dp, Rp = _oracle_run(c, scheme, eps, params)
    return {"reaction": normwise_err(Rp[:, 1], R[:, 1]), "disp": normwise_err(dp.f("DISP"), d.f("DISP")),
            "v": normwise_err(dp.f("V"), d.f("V")), "sigma": normwise_err(dp.f("SIGMA"), d.f("SIGMA")),
            "pl_strain": normwise_err(dp.f("PLSTRAIN"), d.f("PLSTRAIN"))}


def _assert_plane(g):
    x, v, a = g.get("x"), g.get("v"), g.get("a")
    assert np.all(x[:, 2] == 0.0) and np.all(v[:, 2] == 0.0) and np.all(a[:, 2] == 0.0)


@pytest.fixture(scope="module")
def plate():
    c = cases.plate2d_small(steps=300, v_top=-5.0)
    d, R = _oracle_run(c)
    assert d.f("PLSTRAIN").max() > 0
    return c, d, R, {1e-15: _sens(c, d, R, 1e-15), 1e-7: _sens(c, d, R, 1e-7)}


@pytest.mark.parametrize("precision,tol,eps", [(capi.FP64, 1e-8, 1e-15), (capi.MIXED, 1e-4, 1e-7)])
def test_2d_run_matches_oracle(plate, precision, tol, eps):
    c, d, R, sens_all = plate
    sens = sens_all[eps]
    _log(f"2D plate: n={c.n} steps={c.steps} precision={precision}")
    g = cases.make_solver(c, precision=precision)
    g.step(c.steps)
    g.sync()
    errs = _fields(g, d, g.reaction_history(c.top_zone), R)
    _log("2D errors", errs, "oracle sensitivity", sens)
    for k, e in errs.items():
        assert e < max(tol, 10 * sens[k]), k
    _assert_plane(g)
    # the search grid has one z layer (Source/Neighbour.cpp:56-75)
    ref, srch = g.cell_grids()
    assert ref[2] == 1 and srch[2] == 1
    g.close()


def test_2d_single_passes_fp64():
    c = cases.plate2d_small(steps=1)
    rng = np.random.default_rng(3)
    v = np.zeros((c.n, 3))
    v[:, :2] = rng.normal(0.0, 1.0, size=(c.n, 2))
    s = rng.normal(0.0, 1e8, size=(c.n, 3, 3))
    sig = 0.5 * (s + np.transpose(s, (0, 2, 1)))  # with xz, yz parts: the pair forces still lose z
    g = cases.make_solver(c)
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
    a = g.get("a")
    assert np.all(a[:, 2] == 0.0)
    assert normwise_err(a, d.f("A")) < 1e-10
    g.eval("density")
    d.calc_dens()
    assert normwise_err(g.get("ddensity"), d.f("DRHO")) < 1e-10
    g.eval("rates")
    d.calc_rates()
    assert normwise_err(g.get("strain_rate"), d.f("SR")) < 1e-10
    assert normwise_err(g.get("rot_rate"), d.f("RR")) < 1e-10
    g.close()


def test_2d_kickdrift_matches_oracle():
    c = cases.plate2d_small(steps=150, v_top=-5.0)
    d, R = _oracle_run(c, "kickdrift")
    sens = _sens(c, d, R, 1e-15, "kickdrift")
    g = cases.make_solver(c)
    g.step(c.steps, capi.KICKDRIFT)
    g.sync()
    errs = _fields(g, d, g.reaction_history(c.top_zone), R)
    _log("2D KickDrift errors", errs, "sensitivity", sens)
    for k, e in errs.items():
        assert e < max(1e-8, 10 * sens[k]), k
    _assert_plane(g)
    g.close()


def test_2d_thermal_convection_matches_oracle():
    """Conduction, plastic heating and convection on the top row, whose
    dS2 is (m/rho)^(1/3) in plane strain (Source/Thermal.cpp:354-366)."""
    c = cases.plate2d_small(steps=150, v_top=-5.0)
    m = dict(c.materials[0])
    m.update(k_T=200.0, cp_T=900.0)
    c.materials[0] = m
    c.thermal = True
    c.pl_heating = True
    c.tbcs = [dict(zone=2, T=500.0)]
    c.thbc = (c.ids == 3).astype(np.int32)
    c.hconv = np.where(c.ids == 3, 1.0e4, 0.0)
    c.tinf = np.full(c.n, 293.0)
    d, R = _oracle_run(c)
    dp, _ = _oracle_run(c, eps=1e-15)
    g = cases.make_solver(c)
    g.step(c.steps)
    g.sync()
    T_o = d.f("T")
    assert T_o.max() > 400.0
    e_T = normwise_err(g.get("T") - 300.0, T_o - 300.0)
    s_T = normwise_err(dp.f("T") - 300.0, T_o - 300.0)
    e_c = normwise_err(g.get("q_conv"), d.f("QCONV"))
    _log("2D thermal errors T", e_T, "q_conv", e_c, "sensitivity T", s_T)
    assert np.abs(d.f("QCONV")).max() > 0
    assert e_T < max(1e-8, 10 * s_T) and e_c < max(1e-8, 10 * s_T)
    g.close()


def test_2d_gradient_correction_locking_matches_oracle():
    c = cases.plate2d_small(steps=100, v_top=-5.0)
    params = {"nonlock_sum": 0, "grad_corr": 1}
    d, R = _oracle_run(c, params=params)
    sens = _sens(c, d, R, 1e-15, params=params)
    g = cases.make_solver(c, grad_corr=True, locking=True)
    g.step(c.steps)
    g.sync()
    errs = _fields(g, d, g.reaction_history(c.top_zone), R)
    _log("2D gradient-correction errors", errs, "sensitivity", sens)
    for k, e in errs.items():
        assert e < max(1e-8, 10 * sens[k]), k
    L = g.get("gradcorr").reshape(-1, 3, 3)
    assert np.all(L[:, 2, 2] == 1.0)
    _assert_plane(g)
    g.close()


def test_2d_z_velocity_reset_each_step():
    c = cases.plate2d_small(steps=1, v_top=-5.0)
    c.v = np.zeros((c.n, 3))
    c.v[:, 2] = 1.0
    g = cases.make_solver(c)
    assert np.all(g.get("v")[:, 2] == 1.0)
    g.step(1)
    g.sync()
    _assert_plane(g)
    g.close()


def test_2d_refusals():
    c = cases.plate2d_small(steps=1)
    s = capi.Solver(c.n)
    with pytest.raises(capi.WFError) as e:
        s.set_dimension(4)
    assert e.value.code == capi.EINVAL
    s.close()
    # particles off the plane
    c2 = cases.plate2d_small(steps=1)
    c2.x = c2.x.copy()
    c2.x[0, 2] = 1e-4
    with pytest.raises(capi.WFError, match="same z"):
        cases.make_solver(c2)
    # rigid contact
    c3 = cases.plate2d_small(steps=1)
    c3.meshes = [dict(zone=10, axis=1, flip=1, p1=(0.0, 0.061, -1e-3), p2=(0.03, 0.061, 1e-3), dens=4, hfac=1.1)]
    c3.contact = (0.2, 0.2, 0.0)
    with pytest.raises(capi.WFError, match="contact") as e:
        cases.make_solver(c3)
    assert e.value.code == capi.EUNSUPPORTED
    # after init
    g = cases.make_solver(cases.plate2d_small(steps=1))
    with pytest.raises(capi.WFError) as e:
        g.set_dimension(3)
    assert e.value.code == capi.ESTATE
    g.close()


def test_2d_damage_surface_matches_oracle():
    """Damage refreshes the free surface every step (Source/SolverLeapfrog.
    cpp:204-206); in 2D a surface particle has at most 12 neighbours (Source/
    Domain.cpp:1474), so on the plain lattice exactly the outer ring takes
    ID 1.  FP64 with the Locking sums, against the oracle."""
    c = cases.plate2d_small(steps=20, v_top=-5.0, jitter=0.0)
    c.damage = {0: (10.0, 0.0, 0.0, 0.0, 0.0)}
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    for _ in range(c.steps):
        d.step()
    g = cases.make_solver(c)
    g.step(c.steps)
    g.sync()
    ids = g.get_int("id")
    assert np.array_equal(ids, d.i("ID")[:c.n])
    x = c.x
    Lx, Ly = c.meta["Lx"], c.meta["Ly"]
    ring = (x[:, 0] < 1e-9) | (x[:, 0] > Lx - 1e-9) | (x[:, 1] < 1e-9) | (x[:, 1] > Ly - 1e-9)
    assert np.all(ids[ring] == 1) and not np.any(ids[~ring] == 1)
    assert normwise_err(g.get("disp"), d.f("DISP")) < 1e-8
    _assert_plane(g)
    g.close()
