"""Fraser's, Seo's and LS-Dyna's contact laws on the GPU against the CPU
oracle (Source/Contact.cpp:223-770, Source/Contact_LS.cpp:15-292,
dispatched by Source/SolverLeapfrog.cpp:234-239).

Under Leapfrog these laws do not feed their force back into the motion
(only Fraser's Fr_Bound friction changes a), so the runs are not chaotic:
a spinning billet whose ends move towards two rigid planes (the top plane
driven down) for 100 steps, with every particle's contact force of the
last step, the displacement and the sums compared.  Bars: FP64 within
1e-8 normwise (contact force, displacement, contact_force_sum history);
mixed precision within 1e-4; contact pairs and ids exact in FP64.  Two
runs add the adaptive step with the contact criterion (Seo's and Fraser's
min_force_ts; the particles move, so every dt_fext is positive): the step
size history within the same bars."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

STEPS = 100


def _case(law, bound=False, linear=None, auto_ts=0):
    c = cases.contact_small(steps=STEPS, v_punch=-5.0, mu=0.2)
    x = c.x
    zmid, half = 0.5 * (x[:, 2].min() + x[:, 2].max()), 0.5 * (x[:, 2].max() - x[:, 2].min())
    w = 100.0
    c.v = np.stack([-w * x[:, 1], w * x[:, 0], 2.0 * (x[:, 2] - zmid) / half], axis=1)
    c.contact_law = (law, 0.7, bound)
    c.friction_linear = linear
    c.auto_ts = auto_ts
    return c


def _oracle(c):
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    S = []
    for _ in range(STEPS):
        d.step()
        S.append((d.param("contact_force_sum"), d.param("contact_reaction_sum"), d.param("ext_forces_work"),
                  d.param("deltat")))
    return d, np.array(S)


CASES = {"fraser": (capi.FRASER, False, None), "fraser_bound": (capi.FRASER, True, None),
         "seo": (capi.SEO, False, None), "lsdyna": (capi.LSDYNA, False, None),
         "lsdyna_linear_mu": (capi.LSDYNA, False, (2e-4, 0.1)),
         # adaptive step with the contact criterion (auto_ts bits 0 and 2)
         # (LS-Dyna's penalty 2 m PFAC / dt^2 makes that criterion shrink the
         # step towards 0, as with Wang's law; Fraser's and Seo's stiffness
         # does not depend on dt)
         "seo_contact_step": (capi.SEO, False, None, 5), "fraser_contact_step": (capi.FRASER, False, None, 5)}


@pytest.fixture(scope="module")
def oracle_runs():
    out = {}
    for name, args in CASES.items():
        c = _case(*args)
        out[name] = (c,) + _oracle(c)
    return out


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_contact_law_matches_oracle(oracle_runs, name, precision):
    c, d, S = oracle_runs[name]
    fp64 = precision == capi.FP64
    print(f"contact law {name}: n={c.n} steps={STEPS} precision={precision}", flush=True)
    cf_o = d.f("CONTFORCE")[:c.n]
    assert np.abs(cf_o).max() > 0 and np.abs(S[:, 0]).max() > 0
    g = cases.make_solver(c, precision=precision)
    Sg = []
    for _ in range(STEPS):
        g.step(1)
        sc = g.scalars()
        Sg.append((sc["contact_force_sum"], sc["contact_reaction_sum"], sc["ext_forces_work"], sc["dt"]))
    g.sync()
    Sg = np.array(Sg)
    sc = g.scalars()
    errs = {"contforce": normwise_err(g.get("contforce"), cf_o),
            "disp": normwise_err(g.get("disp"), d.f("DISP")[:c.n]),
            "cf_sum": normwise_err(Sg[:, 0], S[:, 0])}
    if c.contact_law[0] == capi.LSDYNA:
        errs["reaction_sum"] = normwise_err(Sg[:, 1], S[:, 1])
        errs["ext_work"] = normwise_err(Sg[:, 2], S[:, 2])
    if c.auto_ts:
        errs["dt"] = normwise_err(Sg[:, 3], S[:, 3])
        assert S[:, 3].min() < 0.5 * c.dt  # the contact criterion acts
    print(f"errors {errs} pairs {sc['contact_pairs']} / {int(d.param('ncont'))}", flush=True)
    bar = 1e-8 if fp64 else 1e-4
    for k, e in errs.items():
        assert e < bar, k
    if fp64:
        assert sc["contact_pairs"] == int(d.param("ncont"))
        assert np.array_equal(g.get_int("id"), d.i("ID")[:c.n])
    g.close()
