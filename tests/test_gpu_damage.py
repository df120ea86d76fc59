"""Johnson-Cook damage on the GPU against the CPU oracle (Domain::CalcDamage,
Source/Damage.cpp:9-46; JohnsonCookDamage::CalcFractureStrain,
Source/Material.cpp:76-92; the pair weights of the Locking passes,
Source/InteractionAlt.cpp:54-60,558-572; the damage schedule and free
surface of Source/SolverLeapfrog.cpp:160-206,315).

A 4.5k-particle JC aluminium billet upset at -20 m/s with the scene
driver's damage settings (Locking sums, FP64), 200 Leapfrog steps.  Bars:
damage, displacement, plastic strain and the top zone's reaction within
max(1e-8, 10x the oracle's own sensitivity to a 1e-15 relative perturbation
of the initial positions) normwise; ids (the per-step free surface) exact
while the run is below that sensitivity at step 100; no neighbour search
after the first step.  The refusals: mixed precision, the Nishimura sums,
KickDrift steps and a non-JC material."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

STEPS = 200
MID = 100
D = (0.2, 0.3, -2.0, 0.01, 0.4)


def _oracle(c, eps=0.0):
    d = oracle_from_case(c, threads=8)
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:c.n] = c.x * (1 + eps * rng.uniform(-1, 1, size=c.x.shape))
    d.solve_init(c.dt)
    R, mid = [], None
    for s in range(STEPS):
        d.step()
        R.append(d.reaction(3))
        if s + 1 == MID:
            mid = (d.f("DAMD")[:c.n].copy(), d.i("ID")[:c.n].copy())
    return d, np.array(R), mid


def _fields(dam, disp, pl, R):
    return {"dam_D": dam, "disp": disp, "pl_strain": pl, "reaction": R}


@pytest.fixture(scope="module")
def damage_run():
    c = cases.damage_small(D, steps=STEPS)
    d, R, mid = _oracle(c)
    ref = _fields(d.f("DAMD")[:c.n].copy(), d.f("DISP")[:c.n].copy(), d.f("PLSTRAIN")[:c.n].copy(), R)
    dp, Rp, _ = _oracle(c, 1e-15)
    pert = _fields(dp.f("DAMD")[:c.n], dp.f("DISP")[:c.n], dp.f("PLSTRAIN")[:c.n], Rp)
    sens = {k: normwise_err(pert[k], ref[k]) for k in ref}
    return c, d, ref, mid, sens


def test_damage_run_matches_oracle(damage_run):
    c, d, ref, mid, sens = damage_run
    print(f"damage run: n={c.n} steps={STEPS} D={D}", flush=True)
    assert ref["dam_D"].max() > 0.05 and np.abs(ref["pl_strain"]).max() > 1e-3
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(MID)
    g.sync()
    dam_mid, id_mid = g.get("dam_D"), g.get_int("id")
    e_mid = normwise_err(dam_mid, mid[0])
    g.step(STEPS - MID)
    g.sync()
    got = _fields(g.get("dam_D"), g.get("disp"), g.get("pl_strain"), g.reaction_history(3))
    errs = {k: normwise_err(got[k], ref[k]) for k in ref}
    sc = g.scalars()
    print("damage errors", errs, "oracle sensitivity", sens, "mid damage", e_mid, "searches", sc["nb_searches"],
          flush=True)
    for k, e in errs.items():
        assert e < max(1e-8, 10 * sens[k]), k
    if e_mid < 1e-8:
        assert np.array_equal(id_mid, mid[1])
    assert sc["nb_searches"] == 1 == d.param("nb_searches")
    g.close()


def test_damage_refusals():
    c = cases.damage_small(D, steps=1)
    for kw in (dict(precision=capi.MIXED), dict(locking=False)):
        s = capi.Solver(c.n, precision=kw.get("precision", capi.FP64), locking=kw.get("locking", True))
        for idx, m in c.materials.items():
            s.set_material(idx, **m)
        for f in ("x", "h", "mass", "density", "sigmay"):
            s.set(f, getattr(c, f))
        s.set_int("id", c.ids)
        s.set_int("material", c.material)
        s.set_material_damage(0, 1, D)
        with pytest.raises(capi.WFError) as e:
            s.init(c.dt)
        assert e.value.code == capi.EUNSUPPORTED
        s.close()
    g = cases.make_solver(c, precision=capi.FP64)
    with pytest.raises(capi.WFError) as e:
        g.step(1, scheme=capi.KICKDRIFT)
    assert e.value.code == capi.EUNSUPPORTED
    g.close()
