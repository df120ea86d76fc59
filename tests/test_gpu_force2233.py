"""The fused pair pass CalcForce2233 (Source/Interaction.cpp:25-333, over
every pair as Domain::LastComputeAcceleration applies it,
Source/Domain.cpp:1700-1709) on the GPU against the CPU oracle: on a
jittered billet with a linear velocity field plus noise and a varying
stress, with and without the kernel-gradient correction, the acceleration,
density rate, strain and rotation rates, ZWab, VXSPH and the Shepard sum
within 1e-10 normwise (FP64: the GPU gathers each particle's terms in row
order, the reference adds them pair by pair).  A uniform velocity gives
VXSPH = 0 exactly."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

XSPH = 0.5


def _state(c, seed=4):
    rng = np.random.default_rng(seed)
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    v = c.x @ A.T + rng.normal(0, 0.05, size=(c.n, 3))
    s = rng.normal(0, 1e8, size=(c.n, 3, 3))
    return v, 0.5 * (s + np.transpose(s, (0, 2, 1)))


@pytest.mark.parametrize("grad_corr", [False, True])
def test_force2233_matches_oracle(grad_corr):
    c = cases.small_jc(steps=1)
    v, sig = _state(c)
    print(f"force2233: n={c.n} grad_corr={grad_corr}", flush=True)
    d = oracle_from_case(c, threads=8)
    d.set_param("nonlock_sum", 0)
    d.set_param("grad_corr", int(grad_corr))
    d.f("V")[:] = v
    d.f("SIGMA")[:] = sig
    d.neighbour_search()
    if grad_corr:
        d.calc_grad_corr()
    d.start_acceleration()
    d.calc_force2233(xsph=XSPH, shepard=True)
    ref = {"a": d.f("A"), "ddensity": d.f("DRHO"), "strain_rate": d.f("SR").reshape(-1, 9),
           "rot_rate": d.f("RR").reshape(-1, 9), "zwab": d.f("ZWAB"), "vxsph": d.f("VXSPH"),
           "sumden": d.f("SUMDEN")}
    assert all(np.abs(r).max() > 0 for r in ref.values())
    g = cases.make_solver(c, precision=capi.FP64, grad_corr=grad_corr, locking=True)
    g.set("v", v)
    g.set("sigma", sig.reshape(c.n, 9))
    g.rebuild_neighbours()
    g.calc_force2233(xsph=XSPH, shepard=True)
    got = {"a": g.get("a"), "ddensity": g.get("ddensity"), "strain_rate": g.get("strain_rate").reshape(-1, 9),
           "rot_rate": g.get("rot_rate").reshape(-1, 9), "zwab": g.get("zwab"), "vxsph": g.get("vxsph"),
           "sumden": g.get("sumden")}
    errs = {k: normwise_err(got[k], ref[k]) for k in ref}
    print("force2233 errors", errs, flush=True)
    for k, e in errs.items():
        assert e < 1e-10, k
    g.set("v", np.broadcast_to([0.3, -1.2, 2.0], (c.n, 3)).copy())
    g.calc_force2233(xsph=XSPH, shepard=False)
    assert np.all(g.get("vxsph") == 0) and np.all(g.get("sumden") == 0)
    g.close()
