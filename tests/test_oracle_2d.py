"""Plane strain (Domain::Dimension 2) in the oracle, pinned by closed forms
and lattice sums that do not run the oracle's own formulas:

- the 2D cubic spline: W(0) = 10/(7 pi h^2), unit integral over the plane,
  the gradient factor equal to dW/dr / r (Source/Functions.cpp:25-87);
- the pair passes on a square lattice under a linear velocity field: the
  interior strain rate, rotation rate and density rate follow from one 2D
  lattice moment computed here from the kernel and the offsets alone, with
  no z components (plane strain);
- StartAcceleration's reset of the z velocity (Source/Domain.cpp:1592);
- the free-surface threshold of 12 neighbours in 2D (Source/Domain.cpp:
  1472-1474): on a plain lattice exactly the outer ring is surface, while
  the second ring (17 neighbours, |n| above h/4) would pass the 3D bound;
- the kernel-gradient correction's zz entry (Source/Domain.cpp:1911-1915);
- contact and the tensile-instability term are refused in 2D."""
import math

import numpy as np
import pytest

import pyoracle as po
from oracle_case import oracle_from_case
from weldform_amd import cases


def _plate(**kw):
    kw.setdefault("jitter", 0.0)
    return cases.plate2d_small(**kw)


def test_2d_kernel_closed_form_and_normalisation():
    h = 1.3e-3
    assert po.kernel(0.0, h, 2) == pytest.approx(10.0 / (7.0 * math.pi * h * h), rel=1e-15)
    # the 3D constant is a different one
    assert po.kernel(0.0, h, 3) == pytest.approx(1.0 / (math.pi * h ** 3), rel=1e-15)
    r = np.linspace(0.0, 2.0 * h, 200001)
    w = np.array([po.kernel(q, h, 2) for q in r[::100] / h])
    rr = r[::100]
    integral = np.trapezoid(w * 2.0 * math.pi * rr, rr)
    assert integral == pytest.approx(1.0, abs=2e-6)
    for q in (1.0, 2.0):
        assert po.kernel(q - 1e-12, h, 2) == pytest.approx(po.kernel(q + 1e-12, h, 2), abs=1e-6 / h ** 2)


def test_2d_grad_kernel_is_dW_dr_over_r():
    h = 0.9e-3
    for q in (0.2, 0.7, 1.3, 1.8):
        r, e = q * h, 1e-7 * h
        dwdr = (po.kernel((r + e) / h, h, 2) - po.kernel((r - e) / h, h, 2)) / (2 * e)
        assert po.grad_kernel(q, h, 2) == pytest.approx(dwdr / r, rel=1e-6)


def _lattice_moment_2d(c):
    """mu with sum_j V_j GK_ij x_ij (x) x_ij = -mu I_2 over the square
    lattice's interior neighbourhood (every offset within 2h), from the
    kernel and the offsets alone."""
    dx, h = c.meta["dx"], c.meta["h"]
    V = c.mass[0] / c.density[0]
    r = np.arange(-3, 4)
    k = np.stack(np.meshgrid(r, r, indexing="ij"), -1).reshape(-1, 2)
    k2 = (k * k).sum(1)
    k2 = k2[(k2 > 0) & (np.sqrt(k2) * dx / h <= 2.0)]
    assert len(k2) == 20
    return -V / 2.0 * sum(po.grad_kernel(math.sqrt(q) * dx / h, h, 2) * q * dx * dx for q in k2)


def _interior(d):
    nb = d.i("NB")
    return nb == nb.max()


def test_2d_linear_velocity_field():
    L = np.array([[1.0, 2.0, 0.0], [0.3, -1.5, 0.0], [0.0, 0.0, 0.0]]) * 100.0
    c = _plate(steps=1)
    mu = _lattice_moment_2d(c)
    assert 0.9 < mu < 1.3
    d = oracle_from_case(c)
    d.f("V")[:] = c.x @ L.T
    d.neighbour_search()
    d.start_acceleration()
    d.calc_rates()
    d.calc_dens()
    m = _interior(d)
    assert m.sum() > 1000
    scale = np.abs(L).max()
    sr = d.f("SR")[m].reshape(-1, 3, 3)
    rr = d.f("RR")[m].reshape(-1, 3, 3)
    assert np.abs(sr - mu * 0.5 * (L + L.T)).max() < 1e-10 * scale
    assert np.abs(rr - mu * 0.5 * (L - L.T)).max() < 1e-10 * scale
    # plane strain: no z rates at all
    assert np.all(d.f("SR").reshape(-1, 3, 3)[:, 2, :] == 0) and np.all(d.f("SR").reshape(-1, 3, 3)[:, :, 2] == 0)
    assert np.abs(d.f("DRHO")[m] + 2700.0 * mu * np.trace(L)).max() < 1e-10 * 2700.0 * scale


def test_2d_step_resets_z_velocity_and_keeps_the_plane():
    c = _plate(steps=3, v_top=-5.0)
    c.v = np.zeros((c.n, 3))
    c.v[:, 2] = 1.0
    d = oracle_from_case(c)
    d.solve_init(c.dt)
    d.step()
    assert np.all(d.f("V").reshape(-1, 3)[:, 2] == 0.0)
    assert np.all(d.f("X").reshape(-1, 3)[:, 2] == 0.0)
    assert np.all(d.f("A").reshape(-1, 3)[:, 2] == 0.0)
    # in 3D the same state keeps its z velocity through the step
    c3 = _plate(steps=3, v_top=-5.0)
    c3.dim = 3
    c3.dom_max, c3.dom_min = (-1e11, -1e11, 5e-3), (1e11, 1e11, -5e-3)  # room for 3D cells along z
    c3.v = c.v.copy()
    d3 = oracle_from_case(c3)
    d3.solve_init(c3.dt)
    d3.step()
    free = c3.ids == 0
    assert np.all(d3.f("V").reshape(-1, 3)[free, 2] == 1.0)


def test_2d_free_surface_is_the_outer_ring():
    c = _plate(steps=2, v_top=-1.0)
    c.damage = {0: (10.0, 0.0, 0.0, 0.0, 0.0)}  # the per-step surface, no damage growth to speak of
    d = oracle_from_case(c)
    d.solve_init(c.dt)
    d.step()
    x = c.x
    dx = c.meta["dx"]
    Lx, Ly = c.meta["Lx"], c.meta["Ly"]
    tol = 1e-9
    ring1 = (x[:, 0] < tol) | (x[:, 0] > Lx - tol) | (x[:, 1] < tol) | (x[:, 1] > Ly - tol)
    ring2 = ~ring1 & ((x[:, 0] < dx + tol) | (x[:, 0] > Lx - dx - tol) | (x[:, 1] < dx + tol) |
                      (x[:, 1] > Ly - dx - tol))
    ids = d.i("ID")
    assert np.all(ids[ring1] == 1)
    assert not np.any(ids[~ring1] == 1)
    # the second ring has 17 neighbours and |n| above h/4: only the 2D bound
    # (12) keeps it off the surface
    nb = d.i("NB")
    mid = ring2 & (x[:, 0] > 5 * dx) & (x[:, 0] < Lx - 5 * dx)
    assert mid.sum() > 10 and np.all(nb[mid] == 17)


def test_2d_gradient_correction_zz_entry():
    c = _plate(steps=1)
    mu = _lattice_moment_2d(c)
    d = oracle_from_case(c)
    d.neighbour_search()
    d.calc_grad_corr()
    m = _interior(d)
    L = d.f("GCM")[m].reshape(-1, 3, 3)
    assert np.all(L[:, 2, 2] == 1.0)
    assert np.all(L[:, 0, 2] == 0.0) and np.all(L[:, 1, 2] == 0.0)
    assert np.abs(L[:, 0, 0] - 1.0 / mu).max() < 1e-9
    assert np.abs(L[:, 0, 1]).max() < 1e-9


def test_2d_refuses_contact_and_tensile_instability():
    c = _plate(steps=1)
    c.materials[0]["ti"] = 0.3
    d = oracle_from_case(c)
    with pytest.raises(RuntimeError, match="tensile"):
        d.solve_init(c.dt)
    c = _plate(steps=1)
    c.meshes = [dict(zone=10, axis=1, flip=1, p1=(0.0, 0.061, -1e-3), p2=(0.03, 0.061, 1e-3), dens=4, hfac=1.1)]
    c.contact = (0.2, 0.2, 0.0)
    d = oracle_from_case(c)
    with pytest.raises(RuntimeError, match="contact"):
        d.solve_init(c.dt)
    with pytest.raises(RuntimeError):
        po.Domain(4).set_param("Dimension", 4)
