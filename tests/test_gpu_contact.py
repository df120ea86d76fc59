"""Rigid-plane contact on the GPU against the CPU oracle: free-surface
detection, contact pairs, the Wang penalty force with Coulomb friction and
the motion of the planes (Source/Contact_Wang.cpp:15-348,
Source/Contact.cpp:10-132, Source/Domain.cpp:1415-1491, Source/Mesh.cpp).

Bars: FP64 free-surface ids and contact-pair counts exact, displacement,
per-step plane forces and particle contact forces within 1e-8 normwise
(the plane force's tangential part carries the reference's static-friction
term in force units, so the check is per component over the history);
mixed precision within 1e-4 on displacement and the normal force."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def contact_run():
    c = cases.contact_small(steps=200)
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    F = []
    for _ in range(c.steps):
        d.step()
        F.append(np.stack([d.contact_force(0), d.contact_force(1)]))
    return c, d, np.array(F)


@pytest.mark.parametrize("precision,tol", [(capi.FP64, 1e-8), (capi.MIXED, 1e-4)])
def test_contact_run_matches_oracle(contact_run, precision, tol):
    c, d, F = contact_run
    print(f"contact run: n={c.n} steps={c.steps} precision={precision}", flush=True)
    g = cases.make_solver(c, precision=precision)
    g.step(c.steps)
    g.sync()
    nrig = sum(2 * m["dens"] ** 2 for m in c.meshes)
    assert np.abs(F[:, 0, 2]).max() > 0 and np.abs(F[:, 1, 2]).max() > 0  # both planes touched
    Fg = np.stack([g.contact_force_history(0), g.contact_force_history(1)], axis=1)
    sc = g.scalars()
    errs = {
        "disp": normwise_err(g.get("disp"), d.f("DISP")[:c.n]),
        "F_top_z": normwise_err(Fg[:, 0, 2], F[:, 0, 2]),
        "F_bottom_z": normwise_err(Fg[:, 1, 2], F[:, 1, 2]),
        "contforce_z": normwise_err(g.get("contforce")[:, 2], d.f("CONTFORCE")[:c.n, 2]),
        "rigid_x": normwise_err(g.rigid("x", nrig), d.f("X")[c.n:]),
    }
    if precision == capi.FP64:
        errs["F_top_xy"] = normwise_err(Fg[:, 0, :2], F[:, 0, :2])
        errs["F_bottom_xy"] = normwise_err(Fg[:, 1, :2], F[:, 1, :2])
    print("contact errors", errs, "pairs", sc["contact_pairs"], int(d.param("ncont")), flush=True)
    if precision == capi.FP64:
        assert np.array_equal(g.get_int("id"), d.i("ID")[:c.n])
        assert sc["contact_pairs"] == int(d.param("ncont"))
        assert np.array_equal(g.rigid("normal", nrig), d.f("NORMAL")[c.n:])
    for k, e in errs.items():
        assert e < tol, k
