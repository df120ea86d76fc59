"""Adjoint of the FP64 Leapfrog step (wf_adj_*) against central finite
differences of the FP64 forward.

The window starts mid-run (the billet already yielding); the loss is
sum_n (F_n - target_n)^2 over the top zone's reaction (no contact: 30 steps,
a snapshot every 10) or the punch plane's contact force (contact: 10 steps,
a snapshot every 4).  The contact window is short because the penalty
contact is chaotic (a perturbation grows ~2x per step once particles touch
the plane): over 10 steps the finite differences converge (the oracle gives
the same difference quotient for relative steps 1e-4, 1e-5 and 1e-6 to
three digits), over 20 or more they do not.  Contact activation switches
within a few 1e-4 of a parameter, so the contact case takes a 1e-5 relative
step (the oracle's quotients at 1e-5 and 1e-6 agree to 1e-5).
Bars: the replayed forward reproduces the first forward bitwise (loss and
end state); every gradient component within 1e-3 relative of the central
difference with a 1e-4 relative parameter step (1e-5 with contact).
Parameters differ in scale by ten orders, so components are compared as
dL/d ln(theta) = theta dL/dtheta; one whose finite difference is below 1e-6
of the largest is compared in absolute terms."""
import numpy as np
import pytest

from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu

JC_KEYS = ("jc_A", "jc_B", "jc_n", "jc_C", "jc_m")


def _fd_check(g, case, mat0, target, kind, index, friction=None, k=10, rel_step=1e-4):
    loss0, hist0 = g.adj_forward(target, checkpoint_every=k, kind=kind, index=index, component=2)
    disp_end = g.get("disp").copy()
    grad = g.adj_backward()
    # backward leaves the window's end state
    assert np.array_equal(g.get("disp"), disp_end)
    # replay reproduces the forward bitwise
    g.adj_restore()
    loss_again, hist_again = g.adj_forward(target, checkpoint_every=k, kind=kind, index=index, component=2)
    assert loss_again == loss0 and np.array_equal(hist_again, hist0)

    def loss_at(key, value):
        g.adj_restore()
        m = dict(mat0)
        if key in JC_KEYS:
            m[key] = value
            g.set_material(0, **m)
        else:
            g.set_contact(value, value, 0.0)
        loss, _ = g.adj_forward(target, checkpoint_every=k, kind=kind, index=index, component=2)
        g.set_material(0, **mat0)
        if friction is not None:
            g.set_contact(friction, friction, 0.0)
        return loss

    checks = [(k, mat0[k], grad[i]) for i, k in enumerate(JC_KEYS)]
    if friction is not None:
        checks.append(("mu", friction, grad[5] + grad[6]))
    fd = {}
    for key, p0, gk in checks:
        h = rel_step * abs(p0)
        fd[key] = (loss_at(key, p0 + h) - loss_at(key, p0 - h)) / (2 * h)
    scaled = {key: (p0 * fd[key], p0 * gk) for key, p0, gk in checks}
    scale = max(abs(v[0]) for v in scaled.values())
    print(f"adjoint loss {loss0:.6e} theta*dL/dtheta (fd, adjoint)", scaled, flush=True)
    assert scale > 0
    for key, (sf, sg) in scaled.items():
        if abs(sf) > 1e-6 * scale:
            assert abs(sg - sf) <= 1e-3 * abs(sf), (key, sg, sf)
        else:
            assert abs(sg - sf) <= 1e-6 * scale, (key, sg, sf)


def test_adjoint_zone_reaction_matches_fd():
    # JC billet at config-1 size, T0 = 400 K so the thermal-softening
    # exponent m matters, top row driven at -20 m/s
    c = cases.cylinder_case("cyl_upset_4k_jc_adjoint", 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"],
                            cases.AL["nu"], cases.JC_AL, steps=130, jitter=1e-3, seed=7, v_top=-20.0, T0=400.0)
    print(f"adjoint zone case: n={c.n}", flush=True)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(100)
    g.sync()
    assert g.scalars()["max_pl_strain"] > 0
    _, hist = g.adj_forward(np.zeros(30), checkpoint_every=10, kind=capi.ADJ_ZONE, index=c.top_zone, component=2)
    g.adj_restore()
    _fd_check(g, c, c.materials[0], 0.98 * hist, capi.ADJ_ZONE, c.top_zone)


def _al_billet(name, steps, **mat):
    return cases.cylinder_case(name, 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"], cases.AL["nu"],
                               dict(cases.JC_AL, **mat), steps=steps, jitter=1e-3, seed=7, v_top=-20.0, T0=400.0)


@pytest.mark.parametrize("window,n_exp", [(100, 0.32), (400, 1.0)])
def test_adjoint_zone_long_windows(window, n_exp):
    """Five and twenty segments of 20 steps.  With the aluminium's n = 0.32
    the window stays at 100 steps: JC hardening B n pl^(n-1) is unbounded as
    pl -> 0, and the reference's explicit return map (scale-back with the
    previous step's Sy) then has a tangent eigenvalue -H/(3G) beyond -1 for
    freshly yielded particles, so over a few hundred steps the exact
    derivative grows without bound while finite differences saturate.  With
    linear hardening (n = 1) the derivative stays well posed over 400
    steps."""
    c = _al_billet(f"cyl_upset_4k_jc_n{n_exp:g}", 100 + window, jc_n=n_exp)
    print(f"adjoint long window: n={c.n} window={window} jc_n={n_exp}", flush=True)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(100)
    g.sync()
    _, hist = g.adj_forward(np.zeros(window), checkpoint_every=20, kind=capi.ADJ_ZONE, index=c.top_zone, component=2)
    g.adj_restore()
    _fd_check(g, c, c.materials[0], 0.98 * hist, capi.ADJ_ZONE, c.top_zone, k=20, rel_step=1e-6)


def test_adjoint_plane_force_matches_fd():
    c = cases.contact_small(steps=310)
    print(f"adjoint contact case: n={c.n}", flush=True)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(300)
    g.sync()
    assert np.abs(g.contact_force_history(0)[-1]).max() > 0
    assert g.scalars()["max_pl_strain"] > 0
    _, hist = g.adj_forward(np.zeros(10), checkpoint_every=4, kind=capi.ADJ_PLANE, index=0, component=2)
    g.adj_restore()
    _fd_check(g, c, c.materials[0], 0.98 * hist, capi.ADJ_PLANE, 0, friction=c.contact[0], k=4, rel_step=1e-5)


def test_adjoint_rejects_unsupported():
    c = cases.small_jc(steps=10)
    g = cases.make_solver(c, precision=capi.MIXED)
    with pytest.raises(capi.WFError):
        g.adj_forward(np.zeros(5), kind=capi.ADJ_ZONE, index=c.top_zone)


def test_window_loss_autograd():
    """torch.autograd.Function over the window: theta.grad is the adjoint
    gradient, and a small step against it lowers the loss."""
    import torch

    from weldform_amd.calibration import JC_KEYS as KEYS, Window, WindowLoss

    c = cases.cylinder_case("cyl_upset_4k_jc_calib", 7.5e-3, 30e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"],
                            cases.AL["nu"], cases.JC_AL, steps=120, jitter=1e-3, seed=7, v_top=-20.0, T0=400.0)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(100)
    g.sync()
    mat0 = dict(c.materials[0])
    # target: the window's reaction with A and n perturbed by 2 %
    g.set_material(0, **dict(mat0, jc_A=1.02 * mat0["jc_A"], jc_n=1.02 * mat0["jc_n"]))
    _, target = g.adj_forward(np.zeros(20), checkpoint_every=10, kind=capi.ADJ_ZONE, index=c.top_zone)
    g.adj_restore()
    g.set_material(0, **mat0)
    w = Window(g, 0, mat0, target, checkpoint_every=10, kind=capi.ADJ_ZONE, index=c.top_zone)
    theta = torch.tensor([mat0[k] for k in KEYS], dtype=torch.float64, requires_grad=True)
    loss = WindowLoss.apply(theta, w)
    loss.backward()
    grad = theta.grad.numpy().copy()
    assert np.all(np.isfinite(grad)) and np.abs(grad).max() > 0
    # the same gradient from the C-ABI directly
    g.adj_restore()
    g.set_material(0, **mat0)
    g.adj_forward(target, checkpoint_every=10, kind=capi.ADJ_ZONE, index=c.top_zone)
    assert np.array_equal(g.adj_backward()[:5], grad)
    # a step of 1e-3 relative size along -grad (in parameter-scaled units)
    scale = np.array([mat0[k] for k in KEYS])
    step = -grad * scale
    step = 1e-3 * step / np.abs(step).max()
    with torch.no_grad():
        loss2 = WindowLoss.apply(torch.tensor(scale * (1 + step)), w)
    print("window loss", float(loss.detach()), "after a step", float(loss2), flush=True)
    assert float(loss2) < float(loss.detach())
