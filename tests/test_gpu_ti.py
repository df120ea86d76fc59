"""Tensile instability on the GPU against the CPU oracle: StartAcceleration's
TI stress (-TI/rho^2 times the positive part of sigma, Source/Domain.cpp:
1529-1563) and CalcAccel's pair term (W(r)/W(d0))^n (TIR_i + TIR_j)
(Source/InteractionAlt.cpp:104-110), with the JSON driver's defaults
TI = 0.3, TIn = 4, TIInitDist = dx (Source/WeldForm.cpp:912-921).

Bars: one acceleration pass on a random tensile stress field within 1e-10
normwise (FP64) and 2e-5 (mixed); a 200-step run within 1e-8 (FP64) and
1e-4 (mixed) on displacement and reaction history."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


def _ti_case(steps):
    c = cases.small_jc(steps=steps)
    m = dict(c.materials[0])
    m.update(ti=0.3, ti_n=4.0, ti_dist=c.meta["dx"])
    c.materials[0] = m
    return c


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-10), (capi.MIXED, 2e-5)])
def test_ti_accel_pass(precision, tol):
    c = _ti_case(1)
    rng = np.random.default_rng(21)
    s = rng.normal(0, 2e8, size=(c.n, 3, 3))
    sig = 0.5 * (s + np.transpose(s, (0, 2, 1))) + 1e8 * np.eye(3)   # mostly tensile
    d = oracle_from_case(c, threads=8)
    d.f("SIGMA")[:] = sig
    d.neighbour_search()
    d.start_acceleration()
    d.calc_accel()
    a_ti = d.f("A").copy()
    d.f("TI")[:] = 0.0
    d.start_acceleration()
    d.calc_accel()
    a_off = d.f("A").copy()
    g = cases.make_solver(c, precision=precision)
    g.set("sigma", sig.reshape(c.n, 9))
    g.rebuild_neighbours()
    g.eval("accel")
    e = normwise_err(g.get("a"), a_ti)
    e_term = normwise_err(g.get("a") - a_off, a_ti - a_off)
    print(f"TI accel pass precision={precision}: {e:.3e}, TI term alone {e_term:.3e}", flush=True)
    assert normwise_err(a_ti - a_off, a_ti) > 1e-3   # the term matters
    assert e < tol


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-8), (capi.MIXED, 1e-4)])
def test_ti_run_matches_oracle(precision, tol):
    c = _ti_case(200)
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    R = []
    for _ in range(c.steps):
        d.step()
        R.append(d.reaction(c.top_zone).copy())
    R = np.array(R)
    g = cases.make_solver(c, precision=precision)
    g.step(c.steps)
    g.sync()
    errs = {"disp": normwise_err(g.get("disp"), d.f("DISP")),
            "reaction": normwise_err(g.reaction_history(c.top_zone)[:, 2], R[:, 2])}
    print(f"TI run precision={precision}", errs, flush=True)
    for k, e in errs.items():
        assert e < tol, k
