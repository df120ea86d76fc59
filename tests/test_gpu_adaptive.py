"""Adaptive time stepping on the GPU against the CPU oracle: CheckMinTSVel
(smallest neighbour distance, CFL d / (Cs + |v|)), CheckMinTSAccel
(sqrt_h_a sqrt(h / |a|), the later check, so it wins when both are on) and
AdaptiveTimeStep's harmonic blend (Source/Domain.cpp:206-319), evaluated
between the two halves of the kick (Source/SolverLeapfrog.cpp:256-262,
Source/SolverKickDrift.cpp:214-226).  The step size lives on the device, so
the amplitude-driven BC also checks that the device clock advances by the
adapted steps.

Bars: FP64 step sizes within 1e-10 relative at every step, displacement and
reaction history within 1e-8 normwise; mixed precision step sizes within
1e-4 relative, displacement within 1e-4."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

STEPS = 100


def _case(auto_ts):
    # initial dt = h / Cs, above the velocity limit 0.7 dx / Cs, so the
    # velocity criterion cuts it at once; the acceleration criterion (with the
    # reference's sqrt_h_a = 0.0025) cuts it further once the ramp moves the top
    c = cases.cylinder_case("cyl_upset_4k_jc_auto_ts", 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"],
                            cases.AL["nu"], cases.JC_AL, dt_cfl=1.0, steps=STEPS, jitter=1e-3, seed=7,
                            v_top=-20.0, T0=300.0)
    c.amps = [dict(id=1, t=[0.0, 40 * c.dt, 1.0], v=[0.0, 1.0, 1.0])]
    c.bcs[1].update(value_type=1, amp_id=1, amp_factor=1.0)
    c.auto_ts = auto_ts
    return c


_ORACLE = {}


def _oracle(auto_ts, scheme):
    key = (auto_ts, scheme)
    if key not in _ORACLE:
        c = _case(auto_ts)
        d = oracle_from_case(c, threads=8)
        d.solve_init(c.dt)
        dts, R = [], []
        for _ in range(c.steps):
            d.step(scheme)
            dts.append(d.param("deltat"))
            R.append(d.reaction(c.top_zone).copy())
        _ORACLE[key] = (c, d, np.array(dts), np.array(R))
    return _ORACLE[key]


@pytest.mark.parametrize("auto_ts,scheme,precision", [
    (1, "leapfrog", capi.FP64), (3, "leapfrog", capi.FP64), (1, "kickdrift", capi.FP64),
    (1, "leapfrog", capi.MIXED), (3, "leapfrog", capi.MIXED)])
def test_adaptive_step_matches_oracle(auto_ts, scheme, precision):
    c, d, dts, R = _oracle(auto_ts, scheme)
    print(f"adaptive dt: n={c.n} steps={c.steps} auto_ts={auto_ts} scheme={scheme} precision={precision}",
          flush=True)
    assert dts.min() < 0.7 * c.dt  # the step did adapt
    g = cases.make_solver(c, precision=precision)
    sch = capi.LEAPFROG if scheme == "leapfrog" else capi.KICKDRIFT
    got = []
    for _ in range(c.steps):
        g.step(1, sch)
        got.append(g.scalars()["dt"])
    g.sync()
    got = np.array(got)
    sc = g.scalars()
    fp64 = precision == capi.FP64
    dt_err = np.abs(got / dts - 1).max()
    errs = {"disp": normwise_err(g.get("disp"), d.f("DISP")),
            "reaction": normwise_err(g.reaction_history(c.top_zone)[:, 2], R[:, 2])}
    print("dt rel err", dt_err, "time", sc["time"], d.param("Time"), errs, flush=True)
    assert dt_err < (1e-10 if fp64 else 1e-4)
    assert sc["time"] == pytest.approx(d.param("Time"), rel=1e-10 if fp64 else 1e-4)
    assert errs["disp"] < (1e-8 if fp64 else 1e-4)
    if fp64:
        assert errs["reaction"] < 1e-8
