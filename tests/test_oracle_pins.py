"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no automated tests or golden vectors for this path, so
each oracle function is checked against closed forms, invariants, special
cases and brute force, each chosen so that a plausible slip (a dropped term,
a sign, a transposed index, a factor 2) fails it.
"""
import math

import numpy as np
import pytest

import pyoracle as po
from oracle_case import brute_pairs, oracle_from_case, unordered
from weldform_amd import cases

H = 1.2e-3


# ---------------------------------------------------------------- kernel
def test_kernel_closed_form_and_normalisation():
    h = H
    assert po.kernel(0.0, h) == pytest.approx(1.0 / (math.pi * h ** 3), rel=1e-15)
    # continuity at q = 1 and support end at q = 2
    assert po.kernel(1.0 - 1e-12, h) == pytest.approx(po.kernel(1.0, h), rel=1e-9)
    assert po.kernel(2.0, h) == 0.0 and po.kernel(1.999999, h) > 0.0
    # 3D normalisation: 4 pi int_0^{2h} W r^2 dr = 1
    r = np.linspace(0, 2 * h, 200001)
    w = np.array([po.kernel(x / h, h) for x in r])
    integral = 4 * math.pi * np.trapezoid(w * r * r, r)
    assert integral == pytest.approx(1.0, abs=1e-8)


def test_grad_kernel_is_dW_dr_over_r():
    h = H
    for q in [1e-3, 0.3, 0.7, 0.999, 1.0, 1.3, 1.7, 1.99]:
        r = q * h
        eps = 1e-7 * h
        dwdr = (po.kernel((r + eps) / h, h) - po.kernel((r - eps) / h, h)) / (2 * eps)
        # GK * x_ij = grad_i W  =>  GK * r = dW/dr
        assert po.grad_kernel(q, h) * r == pytest.approx(dwdr, rel=1e-6)
    # q == 0 branch: limit of (dW/dr)/r = -3 C / h^2 with C = 1/(pi h^3)
    assert po.grad_kernel(0.0, h) == pytest.approx(-3.0 / (math.pi * h ** 5), rel=1e-15)
    assert po.grad_kernel(2.0, h) == 0.0


# ---------------------------------------------------------------- geometry
def test_cylinder_counts_and_symmetry():
    # counts quoted for the reference's generator
    assert len(po.gen_cylinder([0, 0, 0], 7.5e-3, 30e-3, 0.5e-3)) == 4524
    x = po.gen_cylinder([0, 0, 0], 7.5e-3, 30e-3, 0.5e-3)
    # quadrant symmetric lattice, centres within R - r, 29 rows
    assert np.unique(np.round(x[:, 2], 12)).size == 29
    assert np.max(np.hypot(x[:, 0], x[:, 1])) <= 7.5e-3 - 0.5e-3 + 1e-12
    key = lambda a: set(map(tuple, np.round(a, 12)))  # noqa: E731
    assert key(x) == key(x * np.array([-1, 1, 1])) == key(x * np.array([1, -1, 1]))


# ---------------------------------------------------------------- neighbours
@pytest.mark.parametrize("jitter", [False, True])
def test_neighbour_set_equals_brute_force(jitter):
    c = cases.small_jc(steps=1) if jitter else cases.config1(steps=1)
    d = oracle_from_case(c)
    d.neighbour_search()
    p = d.pairs()
    u = unordered(p)
    # each unordered pair once, never a self pair
    assert np.all(u[:, 0] < u[:, 1])
    assert len(np.unique(u, axis=0)) == len(u)
    assert np.array_equal(u, brute_pairs(c.x, c.h))
    if not jitter:
        assert len(u) == 107490
    # per-particle neighbour counts (SaveNeighbourData) match the pair list
    nb = np.bincount(p.ravel(), minlength=c.n)
    assert np.array_equal(nb, d.i("NB"))


def test_pair_orientation_follows_half_stencil():
    """temp1 is the particle whose cell scans the other (same cell: the higher
    index, since ListGenerate pushes to the list head)."""
    c = cases.config1(steps=1)
    d = oracle_from_case(c)
    d.solve_init(c.dt)
    d.neighbour_search()
    p = d.pairs()
    cs = np.array([d.param(f"CellSize{k}") for k in range(3)])
    bl = np.array([d.param(f"BLPF{k}") for k in range(3)])
    cn = np.array([d.param(f"CellNo{k}") for k in range(3)]).astype(int)
    cc = np.clip(np.floor((c.x - bl) / cs).astype(int), 0, cn - 1)
    a, b = p[:, 0], p[:, 1]
    dd = cc[b] - cc[a]
    same = np.all(dd == 0, axis=1)
    assert np.all(a[same] > b[same])
    half = (dd[:, 2] == 1) | ((dd[:, 2] == 0) & (dd[:, 1] == 1)) | ((dd[:, 2] == 0) & (dd[:, 1] == 0) & (dd[:, 0] == 1))
    assert np.all(half[~same])


# ---------------------------------------------------------------- pair passes
def _interior(c, d):
    nb = d.i("NB")
    return nb == nb.max()


def _lattice_state(L=None, sig=None):
    c = cases.config1(steps=1)
    d = oracle_from_case(c)
    if L is not None:
        d.f("V")[:] = c.x @ np.asarray(L).T
    if sig is not None:
        d.f("SIGMA")[:] = sig(c.x)
    d.neighbour_search()
    d.start_acceleration()
    return c, d


def test_uniform_fields_give_zero_rates():
    c = cases.config1(steps=1)
    d = oracle_from_case(c)
    d.f("V")[:] = np.array([0.3, -1.2, 2.0])
    d.neighbour_search()
    d.start_acceleration()
    d.calc_dens()
    d.calc_rates()
    assert np.all(d.f("DRHO") == 0) and np.all(d.f("SR") == 0) and np.all(d.f("RR") == 0)


def _lattice_moment(c):
    """mu with sum_j V_j GK_ij x_ij (x) x_ij = -mu I on the full cubic-lattice
    neighbourhood, computed from the kernel function and the lattice offsets
    alone (V = m/rho, the interior neighbourhood is every offset within 2h)."""
    dx, h = c.meta["dx"], c.meta["h"]
    V = c.mass[0] / c.density[0]
    r = np.arange(-3, 4)
    k = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    k2 = (k * k).sum(1)
    k2 = k2[(k2 > 0) & (np.sqrt(k2) * dx / h <= 2.0)]
    assert len(k2) == 56
    return -V / 3.0 * sum(po.grad_kernel(math.sqrt(q) * dx / h, h) * q * dx * dx for q in k2)


def test_linear_velocity_field():
    """v = L x on the lattice interior: SR = mu sym(L), Omega = mu skew(L),
    drho/dt = -rho mu tr(L), with one lattice moment mu (close to 1)."""
    L = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    c, d = _lattice_state(L=L)
    mu = _lattice_moment(c)
    assert 0.9 < mu < 1.3
    d.calc_rates()
    d.calc_dens()
    m = _interior(c, d)
    scale = np.abs(L).max()
    assert np.abs(d.f("SR")[m] - mu * 0.5 * (L + L.T)).max() < 1e-10 * scale
    assert np.abs(d.f("RR")[m] - mu * 0.5 * (L - L.T)).max() < 1e-10 * scale
    assert np.abs(d.f("DRHO")[m] + 2700.0 * mu * np.trace(L)).max() < 1e-10 * 2700.0 * scale
    # a rigid rotation has no strain rate on the symmetric lattice interior
    W = np.array([[0, 1.0, 2.0], [-1.0, 0, -0.5], [-2.0, 0.5, 0]]) * 100.0
    c, d = _lattice_state(L=W)
    d.calc_rates()
    m = _interior(c, d)
    assert np.abs(d.f("SR")[m]).max() < 1e-10 * 100.0
    assert np.abs(d.f("RR")[m] - mu * W).max() < 1e-10 * 200.0


def test_accel_uniform_and_linear_pressure():
    s0 = np.array([[1.0, 0.2, -0.3], [0.2, -2.0, 0.5], [-0.3, 0.5, 0.7]]) * 1e8
    c, d = _lattice_state(sig=lambda x: np.broadcast_to(s0, (len(x), 3, 3)).copy())
    d.calc_accel()
    m = _interior(c, d)
    a = d.f("A")
    assert np.abs(a[m]).max() < 1e-9 * np.abs(a).max()
    g = np.array([3.0, -1.0, 2.0]) * 1e10     # pressure gradient, Pa/m
    c, d = _lattice_state(sig=lambda x: -(1e8 + x @ g)[:, None, None] * np.eye(3)[None])
    mu = _lattice_moment(c)
    d.calc_accel()
    m = _interior(c, d)
    # momentum equation: a = div(sigma)/rho = -mu grad(p)/rho on the lattice
    assert np.abs(d.f("A")[m] + mu * g / 2700.0).max() < 1e-8 * np.abs(g).max() / 2700.0


def test_momentum_conservation():
    c = cases.small_jc(steps=1)
    d = oracle_from_case(c)
    rng = np.random.default_rng(4)
    d.f("V")[:] = rng.normal(size=(c.n, 3))
    s = rng.normal(0, 1e8, size=(c.n, 3, 3))
    d.f("SIGMA")[:] = 0.5 * (s + np.transpose(s, (0, 2, 1)))
    d.neighbour_search()
    d.start_acceleration()
    d.calc_accel()
    ma = d.f("A") * c.mass[:, None]
    assert np.abs(ma.sum(axis=0)).max() < 1e-12 * np.abs(ma).sum()


# ---------------------------------------------------------------- stress update
def _one_particle(model, **mat):
    c = cases.config1(steps=1)
    d = po.Domain(1)
    E, nu, rho = 68.9e9, 0.3, 2700.0
    G = E / (2 * (1 + nu))
    d.f("RHO")[:] = rho
    d.f("REFRHO")[:] = rho
    d.f("CS")[:] = math.sqrt(E / (3 * (1 - 2 * nu)) / rho)
    d.f("G")[:] = G
    d.f("MASS")[:] = 1.0
    d.i("MODEL")[:] = model
    m = dict(model=model, E=E, nu=nu)
    m.update(mat)
    d.set_material(0, **m)
    return d, G, E, c


def test_elastic_increment_closed_form():
    d, G, E, _ = _one_particle(po.BILINEAR)
    d.f("SIGMAY")[:] = 1e12
    e = 2.0e3
    d.f("SR")[0] = np.diag([e, -0.3 * e, 0.1 * e])
    d.f("SR")[0, 0, 1] = d.f("SR")[0, 1, 0] = 0.4 * e
    dt = 1e-7
    sr = d.f("SR")[0].copy()
    d.stress_all(dt)
    dev = sr - np.trace(sr) / 3 * np.eye(3)
    assert np.allclose(d.f("SHEAR")[0], 2 * G * dt * dev, rtol=1e-14, atol=0)
    # effective strain rate of the reference (Particle.cpp:983-987) = sqrt(3/2 dev:dev)
    assert d.f("EFFSR")[0] == pytest.approx(math.sqrt(1.5 * np.sum(dev * dev)), rel=1e-12)


def test_jaumann_rotation_closed_form():
    d, G, E, _ = _one_particle(po.BILINEAR)
    d.f("SIGMAY")[:] = 1e12
    s, w, dt = 1e8, 5e3, 1e-7
    d.f("SHEAR")[0] = np.diag([s, -s, 0.0])
    d.f("RR")[0, 0, 1], d.f("RR")[0, 1, 0] = w, -w
    d.stress_all(dt)
    S = d.f("SHEAR")[0]
    # S W^T + W S for S = diag(s,-s,0), W_01 = w: off-diagonal -2 s w
    assert S[0, 1] == pytest.approx(-2 * s * w * dt, rel=1e-12)
    assert S[1, 0] == pytest.approx(-2 * s * w * dt, rel=1e-12)
    assert S[0, 0] == s and S[1, 1] == -s


def test_radial_return_bilinear_closed_form():
    Et = 1.5e9
    d, G, E, _ = _one_particle(po.BILINEAR)
    Ep = E * Et / (E - Et)
    Sy = 300e6
    d.f("SIGMAY")[:] = Sy
    d.f("EP")[:] = Ep
    e, dt = 5e4, 1e-6      # one large uniaxial deviatoric increment
    d.f("SR")[0] = np.diag([e, -0.5 * e, -0.5 * e])
    d.stress_all(dt)
    S = d.f("SHEAR")[0]
    sig_tr = 2 * G * dt * e * 1.5           # sqrt(3 J2) of the trial deviator
    assert math.sqrt(1.5 * np.sum(S * S)) == pytest.approx(Sy, rel=1e-12)   # scaled to the old Sy
    dep = (sig_tr - Sy) / (3 * G + Ep)
    assert d.f("PLSTRAIN")[0] == pytest.approx(dep, rel=1e-12)
    assert d.f("SIGMAY")[0] == pytest.approx(Sy + dep * Ep, rel=1e-12)
    # plastic strain increment is deviatoric and coaxial with S (Prandtl-Reuss)
    epl = d.f("EPL")[0]
    assert abs(np.trace(epl)) < 1e-12 * np.abs(epl).max()
    assert np.allclose(epl / np.abs(epl).max(), S / np.abs(S).max(), atol=1e-12)


def test_johnson_cook_and_hollomon_yield_closed_form():
    A, B, n, C, eps0, m, Tm, Tt = 175e6, 380e6, 0.32, 0.0015, 1.0, 1.0, 775.0, 300.0
    d, G, E, _ = _one_particle(po.JOHNSON_COOK, A=A, B=B, n=n, C=C, eps_0=eps0, jm=m, Tm=Tm, Tt=Tt)
    pl, T = 0.1, 400.0
    d.f("PLSTRAIN")[:] = pl
    d.f("T")[:] = T
    d.f("SIGMAY")[:] = 1e12      # keeps the step elastic; Sy is re-evaluated
    e = 100.0
    d.f("SR")[0] = np.diag([e, -0.5 * e, -0.5 * e])
    d.stress_all(1e-12)
    sr_eq = 1.5 * e              # eff. strain rate of a uniaxial deviatoric rate
    want = (A + B * pl ** n) * (1 + C * math.log(sr_eq / eps0)) * (1 - ((T - Tt) / (Tm - Tt)) ** m)
    assert d.f("SIGMAY")[0] == pytest.approx(want, rel=1e-12)
    K, mh, sy0 = 500e6, 0.2, 150e6
    d, G, E, _ = _one_particle(po.HOLLOMON, K=K, m=mh, sy0=sy0)
    d.f("PLSTRAIN")[:] = 0.05
    d.f("SIGMAY")[:] = 1e12
    d.stress_all(1e-12)
    assert d.f("SIGMAY")[0] == pytest.approx(K * (0.05 + sy0 / E) ** mh, rel=1e-12)


# ---------------------------------------------------------------- solver loop
def test_free_particles_drift_and_bc():
    """Particles farther apart than 2h have no pairs: Leapfrog keeps v and
    moves x by v dt per step; a velocity-BC zone follows its amplitude."""
    n = 3
    d = po.Domain(n)
    d.f("X")[:] = np.array([[0, 0, 0], [1.0, 0, 0.5], [0, 1.0, 1.0]])
    d.f("H")[:] = 1e-3
    d.f("MASS")[:] = 1.0
    d.f("RHO")[:] = d.f("REFRHO")[:] = 1000.0
    d.f("SIGMAY")[:] = 1e12
    d.f("V")[:] = np.array([[1.0, 2.0, 3.0], [0, 0, 0], [0, 0, 0]])
    d.i("ID")[:] = [0, 0, 5]
    d.set_material(0, model=po.BILINEAR, E=1e9, nu=0.3)
    d.add_amplitude(1, [0.0, 1e-6, 1.0], [0.0, 2.0, 2.0])
    d.add_velocity_bc(5, [0.0, 0.0, -1.0], value_type=1, amp_id=1, amp_factor=0.5)
    dt = 1e-7
    d.solve_init(dt)
    for k in range(7):
        d.step()
    assert np.allclose(d.f("X")[0], np.array([1.0, 2.0, 3.0]) * 7 * dt, rtol=1e-13)
    # amplitude at t = 6 dt: 2 * 6e-7 / 1e-6 = 1.2 ; velocity = 0.5 * 1.2 * (0,0,-1)
    assert d.f("V")[2] == pytest.approx([0.0, 0.0, -0.6], rel=1e-12)


def _pressure_block(Q):
    """A 14 mm aluminium cylinder (dx = 1 mm, no BC zones) at rest whose
    stress is a fixed pressure field p(x) = p0 + g.x + x.Q x / 2: per-particle
    P0 = p(x), Cs = 0 and G = 0, so the stress update returns sigma = -p I
    every step whatever the density and strain rate.  On the lattice interior
    the acceleration is then a(x) = -mu grad p(x) / rho (the symmetric pair
    sum of a quadratic field leaves the gradient term only), with the lattice
    moment mu of the full neighbourhood.  A step of 1e-8 s keeps the lattice
    distortion of two steps (|grad v| dt ~ 1e-8) below the tolerances.
    Returns the case, the oracle, a(x) of the closed form, mu and the masks
    deep[k] of particles whose neighbourhoods are interior to depth k: a
    quantity of step s that reads neighbours' values of step s - 1 follows
    the closed form one depth further in."""
    c = cases.cylinder_case("pin_pressure", 14e-3, 24e-3, 1.0e-3, cases.AL["rho"], cases.AL["E"], cases.AL["nu"],
                            dict(model=po.BILINEAR, alpha=0.0, beta=0.0, Ep=0.0, sy0=1e12), steps=2)
    c.bcs = []
    c.dt = 1e-8
    d = oracle_from_case(c)
    g = np.array([3.0, -1.0, 2.0]) * 1e10
    ctr = c.x.mean(axis=0)
    xr = c.x - ctr
    p = 1e8 + xr @ g + 0.5 * np.einsum("ni,ij,nj->n", xr, Q, xr)
    n = c.n
    d.f("P0")[:n] = p
    d.f("CS")[:n] = 0.0
    d.f("G")[:n] = 0.0
    d.f("SIGMA")[:n] = -p[:, None, None] * np.eye(3)[None]
    mu = _lattice_moment(c)
    a_cf = -mu * (g + xr @ Q) / c.density[0]
    d.neighbour_search()
    nb = d.i("NB")[:n]
    interior = nb == nb.max()
    P = d.pairs()
    def deeper(m):
        bad = np.zeros(n, dtype=bool)
        np.logical_or.at(bad, P[:, 0], ~m[P[:, 1]])
        np.logical_or.at(bad, P[:, 1], ~m[P[:, 0]])
        return m & ~bad
    deep = [interior]
    for _ in range(3):
        deep.append(deeper(deep[-1]))
    assert deep[3].sum() > 100
    return c, d, a_cf, mu, deep


def test_leapfrog_step_composition_closed_form():
    """SolveDiffUpdateLeapFrog from rest (Source/SolverLeapfrog.cpp:249-301):
    step 1 kicks v -= a dt/2 then v += a (dt_prev + dt)/2, so v_1 = a dt/2;
    step 2 gives v_2 = v_1 + a dt = 3/2 a dt.  The density rate of step 1
    sees v_{1/2} = v_1 at x_0 and the rates see v_1 at x_1, so on the
    interior drho_1 = -rho mu tr(grad v_1) and SR_1 = mu sym(grad v_1);
    step 2's rates see v_2 = 3 v_1."""
    Q = np.array([[2.0, 0.5, -0.3], [0.5, -1.0, 0.8], [-0.3, 0.8, 1.5]]) * 1e12
    c, d, a_cf, mu, deep = _pressure_block(Q)
    dt, n, rho = c.dt, c.n, c.density[0]
    d.solve_init(dt)
    d.step()
    m0, m1, m2, m3 = deep
    scale = np.abs(a_cf[m0]).max()
    assert np.abs(d.f("A")[:n][m0] - a_cf[m0]).max() < 1e-9 * scale
    assert np.abs(d.f("V")[:n][m0] - 0.5 * dt * a_cf[m0]).max() < 1e-9 * scale * dt
    grad_v1 = -mu * Q * (0.5 * dt) / rho
    sr_scale = np.abs(grad_v1).max() * mu
    assert np.abs(d.f("SR")[:n][m1] - mu * grad_v1).max() < 1e-6 * sr_scale
    want = -rho * mu * np.trace(grad_v1)
    assert np.abs(d.f("DRHO")[:n][m1] - want).max() < 1e-6 * abs(want)
    d.step()
    assert np.abs(d.f("V")[:n][m2] - 1.5 * dt * a_cf[m2]).max() < 1e-6 * scale * dt
    assert np.abs(d.f("SR")[:n][m3] - 3.0 * mu * grad_v1).max() < 1e-5 * sr_scale


def test_kickdrift_step_composition_closed_form():
    """SolveDiffUpdateKickDrift from rest (Source/SolverKickDrift.cpp:224-279):
    v += a dt_prev/2, density with that v at x_n, drift, v += a dt/2, rates
    with v_{n+1} at x_{n+1}: after step 1 v_1 = a dt, drho_1 sees a dt/2 and
    SR_1 sees a dt; after step 2 v_2 = 2 a dt."""
    Q = np.array([[2.0, 0.5, -0.3], [0.5, -1.0, 0.8], [-0.3, 0.8, 1.5]]) * 1e12
    c, d, a_cf, mu, deep = _pressure_block(Q)
    dt, n, rho = c.dt, c.n, c.density[0]
    d.solve_init(dt)
    d.step("kickdrift")
    m0, m1, m2, m3 = deep
    scale = np.abs(a_cf[m0]).max()
    assert np.abs(d.f("V")[:n][m0] - dt * a_cf[m0]).max() < 1e-9 * scale * dt
    grad_half = -mu * Q * (0.5 * dt) / rho
    want = -rho * mu * np.trace(grad_half)
    assert np.abs(d.f("DRHO")[:n][m1] - want).max() < 1e-6 * abs(want)
    sr_scale = np.abs(grad_half).max() * mu
    assert np.abs(d.f("SR")[:n][m1] - 2.0 * mu * grad_half).max() < 1e-6 * sr_scale
    d.step("kickdrift")
    assert np.abs(d.f("V")[:n][m2] - 2.0 * dt * a_cf[m2]).max() < 1e-6 * scale * dt


def test_reaction_is_read_before_the_bc_zeroes_acceleration():
    """The reaction observer sums m a over each velocity-BC zone after
    AccelReduction and before the BC zeroes a (Source/SolverLeapfrog.cpp:
    225-228, Source/WeldForm.cpp:51-92).  The pair forces are antisymmetric,
    so the zones' reactions plus m a of every other particle sum to zero."""
    c = cases.small_jc(steps=3)
    c.bcs[1]["value"] = (0.0, 0.0, -20.0)
    d = oracle_from_case(c)
    rng = np.random.default_rng(11)
    d.f("V")[:] = rng.normal(0, 1.0, size=(c.n, 3))
    s = rng.normal(0, 1e8, size=(c.n, 3, 3))
    d.f("SIGMA")[:] = 0.5 * (s + np.transpose(s, (0, 2, 1)))
    d.solve_init(c.dt)
    for _ in range(3):
        d.step()
        ids = d.i("ID")
        free = ~np.isin(ids, [b["zone"] for b in c.bcs])
        ma = d.f("A")[free] * c.mass[free, None]
        R = sum(d.reaction(b["zone"]) for b in c.bcs)
        total = np.abs(d.f("A") * c.mass[:, None]).sum() + np.abs(R).sum()
        assert np.abs(R).max() > 1e-3 * total / c.n
        assert np.abs(R + ma.sum(axis=0)).max() < 1e-11 * total
        # the zone particles' acceleration was zeroed after the observation
        assert np.all(d.f("A")[~free] == 0.0)


def test_neighbour_schedule():
    """Searches run on step 1, on the first step whose start-of-step max
    plastic strain exceeds 1e-6, and every ts_nb_inc = 5 steps after."""
    c = cases.small_jc(steps=120)
    c.bcs[1]["value"] = (0.0, 0.0, -20.0)
    d = oracle_from_case(c, threads=4)
    d.solve_init(c.dt)
    searches, mx = [], []
    for _ in range(c.steps):
        before = d.param("nb_searches")
        d.step()
        mx.append(d.param("max_pl_strain"))
        searches.append(d.param("nb_searches") - before)
    s0 = next(k for k, v in enumerate(mx) if v > 1e-6)
    want = [1 if (k == 0 or (k >= s0 and (k - s0) % 5 == 0)) else 0 for k in range(c.steps)]
    assert searches == want


def test_threads_do_not_change_results():
    c = cases.small_jc(steps=30)
    c.bcs[1]["value"] = (0.0, 0.0, -20.0)
    out = []
    for th in (1, 4):
        d = oracle_from_case(c, threads=th)
        d.solve_init(c.dt)
        for _ in range(c.steps):
            d.step()
        out.append((d.f("X").copy(), d.f("SIGMA").copy(), d.param("int_energy_sum")))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


# ---------------------------------------------------------------- thermal
def test_conduction_uniform_and_quadratic_temperature():
    """CalcTempInc on the lattice interior: a uniform T gives exactly zero; a
    quadratic T = T0 + c x^2 (Laplacian 2c) gives rho cp dT/dt = 2 k c mu,
    with the same lattice moment mu as the velocity gradient (all particles
    share m/rho, so the reference's second-partner volume is uniform)."""
    k, cp = 50.0, 460.0
    c = cases.config1(steps=1)
    d = oracle_from_case(c)
    d.set_param("thermal_solver", 1)
    d.f("KT")[:] = k
    d.f("CPT")[:] = cp
    d.f("T")[:] = 300.0
    d.neighbour_search()
    d.thermal_calcs(0.0)
    assert np.all(d.f("DTDT") == 0.0)
    cq = 4.0e6                               # K/m^2
    d = oracle_from_case(c)
    d.set_param("thermal_solver", 1)
    d.f("KT")[:] = k
    d.f("CPT")[:] = cp
    d.f("T")[:] = 300.0 + cq * c.x[:, 0] ** 2
    d.neighbour_search()
    d.thermal_calcs(0.0)
    mu = _lattice_moment(c)
    m = _interior(c, d)
    want = 2.0 * k * cq * mu / (2700.0 * cp)
    assert np.abs(d.f("DTDT")[m] - want).max() < 1e-9 * want


def _two_particles(thermal=True):
    """Particles 0 and 1 interact; 2 and 3 are far away and only give the
    cell grid of CellInitiate its extent."""
    d = po.Domain(4)
    d.f("X")[:] = np.array([[0.0, 0.0, 0.0], [0.8e-3, 0.3e-3, -0.1e-3], [20e-3, 20e-3, 20e-3],
                            [-20e-3, -20e-3, -20e-3]])
    d.f("H")[:] = 1.2e-3
    d.f("MASS")[:] = [1.0e-3, 2.0e-3, 1.0e-3, 1.0e-3]
    d.f("RHO")[:] = d.f("REFRHO")[:] = [1000.0, 4000.0, 1000.0, 1000.0]
    d.f("SIGMAY")[:] = 1e12
    d.f("KT")[:] = [50.0, 20.0, 50.0, 50.0]
    d.f("CPT")[:] = [460.0, 500.0, 460.0, 460.0]
    d.f("T")[:] = [300.0, 350.0, 300.0, 300.0]
    d.i("FREE")[:] = 1
    d.set_material(0, model=po.BILINEAR, E=1e9, nu=0.3)
    d.set_param("thermal_solver", int(thermal))
    d.neighbour_search()
    return d


def test_conduction_pair_closed_form_and_volume_quirk():
    """Two particles: m = V_second 4 k1 k2/(k1+k2) (T_first - T_second) GK
    goes +m to the first and -m to the second partner, with the SECOND
    particle's m/rho for both (Source/Thermal.cpp:53-219)."""
    d = _two_particles()
    (p1, p2), = d.pairs().tolist()
    d.thermal_calcs(0.0)
    x, T, kT = d.f("X"), d.f("T"), d.f("KT")
    h = 1.2e-3
    r = float(np.linalg.norm(x[p1] - x[p2]))
    V2 = d.f("MASS")[p2] / d.f("RHO")[p2]
    m = V2 * 4.0 * kT[p1] * kT[p2] / (kT[p1] + kT[p2]) * (T[p1] - T[p2]) * po.grad_kernel(r / h, h)
    rc = d.f("RHO") * d.f("CPT")
    assert d.f("DTDT")[p1] == pytest.approx(m / rc[p1], rel=1e-13)
    assert d.f("DTDT")[p2] == pytest.approx(-m / rc[p2], rel=1e-13)
    assert np.all(d.f("DTDT")[2:] == 0.0)
    # heat flows from hot to cold
    hot = int(np.argmax(T[:2]))
    assert d.f("DTDT")[hot] < 0 < d.f("DTDT")[1 - hot]


def test_convection_source_and_thermal_strain_rate():
    """Isolated particle: q_conv = rho h_c (m/rho)^(2/3) (T_inf - T) / m
    (Source/Thermal.cpp:339-389), dT/dt = (q_conv + q_src)/(rho cp), and the
    strain rate gains th_exp dT/dt on its diagonal (ThermalStuct.cpp:7-12)."""
    d = po.Domain(3)
    m, rho, hc, Tinf, T0, qs, cp, ath = 2e-4, 7850.0, 35.0, 293.0, 900.0, 5e7, 460.0, 1.2e-5
    d.f("X")[:] = np.array([[0.0, 0.0, 0.0], [20e-3, 20e-3, 20e-3], [-20e-3, -20e-3, -20e-3]])
    d.f("H")[:] = 1e-3
    d.f("MASS")[:] = m
    d.f("RHO")[:] = d.f("REFRHO")[:] = rho
    d.f("T")[:] = T0
    d.f("CPT")[:] = cp
    d.f("KT")[:] = 40.0
    d.f("HCONV")[:] = hc
    d.f("TINF")[:] = Tinf
    d.f("QSRC")[:] = qs
    d.f("THEXP")[:] = ath
    d.i("THBC")[:] = [1, 0, 0]
    d.f("QSRC")[1:] = 0.0
    d.set_material(0, model=po.BILINEAR, E=1e9, nu=0.3)
    d.set_param("thermal_solver", 1)
    d.neighbour_search()
    d.thermal_calcs(0.0)
    qc = rho * hc * (m / rho) ** (2.0 / 3.0) * (Tinf - T0) / m
    assert d.f("QCONV")[0] == pytest.approx(qc, rel=1e-14)
    dTdt = (qc + qs) / (rho * cp)
    assert d.f("DTDT")[0] == pytest.approx(dTdt, rel=1e-14)
    sr = d.f("SR")[0].reshape(3, 3)
    assert np.allclose(np.diag(sr), ath * dTdt, rtol=1e-14) and np.all(sr[~np.eye(3, dtype=bool)] == 0)
    # the next call first advances T with the stored rate
    d.thermal_calcs(1e-3)
    assert d.f("T")[0] == pytest.approx(T0 + 1e-3 * dTdt, rel=1e-15)


def test_plastic_work_heat_closed_forms():
    """q_pl = sigma : (d eps_pl / dt) with both off-diagonal halves counted:
    uniaxial s * e / dt, pure shear tau * gamma / dt.  ps_energy and
    plastic_work integrate it; a particle whose plastic increment stopped
    keeps its last q_pl (delta_pl_strain is never reset in the reference)."""
    dt, s, e, tau, gam = 1e-7, 3e8, 2e-4, 1.5e8, 6e-4
    d = po.Domain(2)
    d.f("H")[:] = 1e-3
    d.f("MASS")[:] = [1e-3, 2e-3]
    d.f("RHO")[:] = [1000.0, 2000.0]
    d.set_param("pl_heating", 1)
    sig = np.zeros((2, 3, 3))
    inc = np.zeros((2, 3, 3))
    sig[0, 0, 0] = s
    inc[0] = np.diag([e, -e / 2, -e / 2])
    sig[1, 0, 1] = sig[1, 1, 0] = tau
    inc[1, 0, 1] = inc[1, 1, 0] = gam / 2
    d.f("SIGMA")[:] = sig
    d.f("EPLINC")[:] = inc
    d.f("DPLSTRAIN")[:] = 1e-4
    d.plastic_work_heat(dt)
    q = d.f("QPL").copy()
    assert q[0] == pytest.approx(s * e / dt, rel=1e-14)
    assert q[1] == pytest.approx(tau * gam / dt, rel=1e-14)
    assert np.allclose(d.f("PSENERGY"), q * dt, rtol=1e-14)
    assert d.param("plastic_work") == pytest.approx((1e-6 * q[0] + 1e-6 * q[1]) * dt, rel=1e-14)
    d.f("DPLSTRAIN")[:] = 0.0
    d.f("SIGMA")[:] = 0.0
    d.plastic_work_heat(dt)
    assert np.array_equal(d.f("QPL"), q)
    assert np.allclose(d.f("PSENERGY"), 2 * q * dt, rtol=1e-14)


def test_energy_sums():
    """Internal energy: sum of (m/rho) S : eps_dot dt with both off-diagonal
    halves (Source/Domain.cpp:1810-1829); the kinetic-energy pair terms
    (Source/Domain.cpp:1752-1808) are added to one partner and subtracted from
    the other, so their sum is exactly zero."""
    d = _two_particles(thermal=False)
    dt = 2e-7
    d.set_param("deltat", dt)
    S = np.array([[1.0, 0.5, -0.2], [0.5, -0.7, 0.3], [-0.2, 0.3, -0.3]]) * 1e8
    R = np.array([[2.0, 1.0, 0.4], [1.0, -1.0, -0.6], [0.4, -0.6, 0.5]]) * 1e2
    d.f("SHEAR")[:2] = np.stack([S, 2 * S])
    d.f("SR")[:2] = np.stack([R, -R])
    d.f("V")[:2] = [[1.0, 0.0, 0.0], [-2.0, 0.5, 0.0]]
    d.f("P")[:2] = [1e8, 3e8]
    d.energies()
    V = d.f("MASS") / d.f("RHO")
    SR = float(np.sum(S * R))
    assert d.param("int_energy_sum") == pytest.approx((V[0] * SR - V[1] * 2 * SR) * dt, rel=1e-13)
    assert d.param("kin_energy_sum") == 0.0


# ---------------------------------------------------------------- time step
def _rest_block(n=5, dx=1e-3, v=None):
    g = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * dx
    d = po.Domain(len(x))
    E, nu, rho = 68.9e9, 0.3, 2700.0
    K = E / (3 * (1 - 2 * nu))
    d.f("X")[:] = x
    d.f("H")[:] = 1.2 * dx
    d.f("MASS")[:] = rho * dx ** 3
    d.f("RHO")[:] = d.f("REFRHO")[:] = rho
    d.f("CS")[:] = math.sqrt(K / rho)
    d.f("G")[:] = E / (2 * (1 + nu))
    d.f("SIGMAY")[:] = 1e12
    if v is not None:
        d.f("V")[:] = v
    d.set_material(0, model=po.BILINEAR, E=E, nu=nu)
    return d, x, math.sqrt(K / rho)


def test_adaptive_time_step_velocity_criterion():
    """CheckMinTSVel + AdaptiveTimeStep (Source/Domain.cpp:206-299): a
    resting lattice gives dt_min = CFL dx / Cs, taken at once when below
    the initial dt; a particle moving at |v| lowers it to CFL dx/(Cs+|v|);
    Time advances by the new dt."""
    dx, dt0 = 1e-3, 1e-6
    d, x, cs = _rest_block()
    d.set_param("auto_ts", 1)
    d.solve_init(dt0)
    d.step()
    want = 0.7 * dx / cs
    assert d.param("deltat") == pytest.approx(want, rel=1e-13)
    assert d.param("Time") == pytest.approx(want, rel=1e-13)
    v = np.zeros((125, 3))
    v[62] = [30.0, 40.0, 0.0]          # |v| = 50
    d, x, cs = _rest_block(v=v)
    d.set_param("auto_ts", 1)
    d.solve_init(dt0)
    d.step()
    assert d.param("deltat") == pytest.approx(0.7 * dx / (cs + 50.0), rel=1e-12)


def test_adaptive_time_step_harmonic_recovery():
    """Below dt_min the step grows by the harmonic mean 2 dt dt_min /
    (dt + dt_min); with no criterion below the initial dt it returns to it
    the same way."""
    dx = 1e-3
    d, x, cs = _rest_block()
    dmin = 0.7 * dx / cs
    d.set_param("auto_ts", 1)
    d.solve_init(1e-6)
    d.set_param("deltat", 0.25 * dmin)
    d.step()
    assert d.param("deltat") == pytest.approx(2 * 0.25 * dmin * dmin / (1.25 * dmin), rel=1e-13)
    d, x, cs = _rest_block()
    dint = 0.5 * dmin
    d.set_param("auto_ts", 1)
    d.solve_init(dint)
    d.set_param("deltat", 0.1 * dint)
    d.step()
    assert d.param("deltat") == pytest.approx(2 * 0.1 * dint * dint / (1.1 * dint), rel=1e-13)


def test_adaptive_time_step_acceleration_criterion():
    """CheckMinTSAccel: a resting lattice has a = 0, so sqrt_h_a sqrt(h/|a|)
    is infinite and dt stays at its initial value; with both criteria the
    acceleration one, computed last, is the one kept."""
    d, x, cs = _rest_block()
    d.set_param("auto_ts", 3)
    d.solve_init(1e-6)
    d.step()
    assert d.param("deltat") == 1e-6


# ---------------------------------------------------------------- tensile instability
def test_sym3_eig_against_lapack():
    """The Jacobi eigen-decomposition reproduces A and agrees with LAPACK's
    (numpy.linalg.eigh) eigenvalues, including repeated ones."""
    rng = np.random.default_rng(8)
    mats = [rng.normal(size=(3, 3)) * 1e8 for _ in range(40)]
    mats += [np.diag([2.0, 2.0, -1.0]) * 1e8, np.eye(3) * 5e7, np.zeros((3, 3))]
    for M in mats:
        A = 0.5 * (M + M.T)
        w, V = po.sym3_eig(A)
        scale = max(np.abs(A).max(), 1.0)
        assert np.abs(V @ np.diag(w) @ V.T - A).max() < 1e-14 * scale
        assert np.abs(V.T @ V - np.eye(3)).max() < 1e-14
        assert np.allclose(np.sort(w), np.linalg.eigh(A)[0], rtol=0, atol=1e-14 * scale)


def _ti_particles(sig, TI=0.3, rho=2700.0):
    n = len(sig)
    d = po.Domain(n)
    d.f("RHO")[:] = rho
    d.f("SIGMA")[:] = np.asarray(sig).reshape(n, 3, 3)
    d.f("TI")[:] = TI
    d.start_acceleration()
    return d.f("TIR").reshape(n, 3, 3).copy()


def test_tensile_instability_stress_is_the_positive_spectral_part():
    """TIR = -TI/rho^2 sigma_+, sigma_+ the positive part of sigma: closed
    form on diagonal stresses, rotation covariance, and LAPACK's
    eigenvectors on random stresses (StartAcceleration, Source/Domain.cpp:
    1545-1563)."""
    TI, rho = 0.3, 2700.0
    c = TI / rho ** 2
    D = np.diag([3e8, -1e8, 5e7])
    t = _ti_particles([D])[0]
    assert np.allclose(t, np.diag([-c * 3e8, 0.0, -c * 5e7]), rtol=1e-15, atol=0)
    # rotation covariance
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1.0]])
    R = R @ np.array([[1, 0, 0], [0, math.cos(0.4), -math.sin(0.4)], [0, math.sin(0.4), math.cos(0.4)]])
    t = _ti_particles([R @ D @ R.T])[0]
    assert np.abs(t - R @ np.diag([-c * 3e8, 0.0, -c * 5e7]) @ R.T).max() < 1e-14 * c * 3e8
    # random stresses: the positive part from LAPACK's decomposition
    rng = np.random.default_rng(9)
    S = rng.normal(size=(30, 3, 3)) * 2e8
    S = 0.5 * (S + np.transpose(S, (0, 2, 1)))
    t = _ti_particles(S)
    for k in range(30):
        w, V = np.linalg.eigh(S[k])
        want = V @ np.diag(np.where(w > 0, -c * w, 0.0)) @ V.T
        assert np.abs(t[k] - want).max() < 1e-13 * c * np.abs(w).max()
    # compressive stress: no correction
    assert np.all(_ti_particles([-np.eye(3) * 1e8])[0] == 0.0)


def test_tensile_instability_pair_term():
    """Two particles: switching TI on adds m_j GK (W(r)/W(d0))^n (TIR_i + TIR_j)
    x_ij to a_i (CalcAccel, Source/InteractionAlt.cpp:104-126)."""
    h, dx, rho, m = H, 1e-3, 2700.0, 2.7e-6
    x = np.array([[0.0, 0.0, 0.0], [0.8e-3, 0.3e-3, -0.2e-3], [5e-3, 5e-3, 5e-3], [-5e-3, -5e-3, -5e-3]])

    def run(TI):
        d = po.Domain(4)
        d.f("X")[:] = x
        d.f("H")[:] = h
        d.f("MASS")[:] = m
        d.f("RHO")[:] = d.f("REFRHO")[:] = rho
        d.f("CS")[:] = 5000.0
        d.f("TI")[:] = TI
        d.f("TIDIST")[:] = dx
        sig = np.zeros((4, 3, 3))
        sig[0] = np.diag([2e8, -1e8, 1e8])
        sig[1] = np.array([[1e8, 5e7, 0], [5e7, 3e8, 0], [0, 0, -2e8]])
        d.f("SIGMA")[:] = sig
        d.set_material(0, model=po.BILINEAR, E=68.9e9, nu=0.3)
        d.neighbour_search()
        d.start_acceleration()
        d.calc_accel()
        return d.f("A").copy(), d.f("TIR").reshape(4, 3, 3).copy()

    a0, _ = run(0.0)
    a1, tir = run(0.3)
    xij = x[0] - x[1]
    r = np.linalg.norm(xij)
    GK = po.grad_kernel(r / h, h)
    f = (po.kernel(r / h, h) / po.kernel(dx / h, h)) ** 4.0
    want = m * GK * (f * (tir[0] + tir[1])) @ xij
    assert np.abs(tir[0]).max() > 0 and np.abs(tir[1]).max() > 0
    assert np.allclose(a1[0] - a0[0], want, rtol=1e-10, atol=0)
    assert np.allclose(a1[1] - a0[1], -want, rtol=1e-10, atol=0)


# ------------------------------------- Locking sums, kernel-gradient correction
def _jittered(grad_corr=False, nonlock=True, A=None):
    c = cases.small_jc(steps=1)
    d = oracle_from_case(c)
    d.set_param("nonlock_sum", 1 if nonlock else 0)
    d.set_param("grad_corr", 1 if grad_corr else 0)
    if A is not None:
        d.f("V")[:] = c.x @ np.asarray(A).T
    d.f("SIGMA")[:] = np.einsum("n,ab->nab", 1.0 + c.x[:, 2] * 30.0, np.array(
        [[1.0, 0.2, -0.3], [0.2, -2.0, 0.5], [-0.3, 0.5, 0.7]]) * 1e8)
    d.neighbour_search()
    d.start_acceleration()
    if grad_corr:
        d.calc_grad_corr()
    return c, d


def test_locking_sums_equal_nishimura_up_to_rounding():
    """The Locking passes add each pair into both particles in pair order;
    the Nishimura passes gather.  Same sums: acceleration and density rate
    to rounding, rates to the float rounding of the Locking weight m_j/rho_j
    (`float mj_dj`, Source/InteractionAlt.cpp:412)."""
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    out = {}
    for nonlock in (True, False):
        c, d = _jittered(nonlock=nonlock, A=A)
        d.calc_accel()
        d.calc_dens()
        d.calc_rates()
        out[nonlock] = {k: d.f(k).copy() for k in ("A", "DRHO", "SR", "RR")}
    for k, tol in (("A", 1e-12), ("DRHO", 1e-12), ("SR", 1e-6), ("RR", 1e-6)):
        a, b = out[True][k], out[False][k]
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err < tol, (k, err)
    assert not np.array_equal(out[True]["SR"], out[False]["SR"])  # the float weights do change them


def test_grad_corr_matrix_pins():
    """gradCorrM = (-sum_j V_j GK x_ij (x) x_ij)^-1: on the lattice interior
    the sum is mu I (the lattice moment), so L = I / mu; everywhere L is
    symmetric and inverts the sum formed here from the brute-force pair set."""
    c = cases.config1(steps=1)
    d = oracle_from_case(c)
    d.set_param("grad_corr", 1)
    d.neighbour_search()
    d.calc_grad_corr()
    L = d.f("GCM").reshape(-1, 3, 3)
    mu = _lattice_moment(c)
    m = _interior(c, d)
    assert np.abs(L[m] - np.eye(3) / mu).max() < 1e-12
    assert np.abs(L - L.transpose(0, 2, 1)).max() < 1e-9 * np.abs(L).max()
    P = brute_pairs(c.x, c.h)
    xij = c.x[P[:, 0]] - c.x[P[:, 1]]
    h = 0.5 * (c.h[P[:, 0]] + c.h[P[:, 1]])
    r = np.linalg.norm(xij, axis=1)
    GK = np.array([po.grad_kernel(q, hh) for q, hh in zip(r / h, h)])
    V = c.mass / c.density
    dyad = GK[:, None, None] * xij[:, :, None] * xij[:, None, :]
    M = np.zeros((c.n, 3, 3))
    np.add.at(M, P[:, 0], -V[P[:, 1], None, None] * dyad)
    np.add.at(M, P[:, 1], -V[P[:, 0], None, None] * dyad)
    assert np.abs(np.einsum("nab,nbc->nac", L, M) - np.eye(3)).max() < 1e-10


def test_grad_corr_reproduces_linear_fields_everywhere():
    """With the correction (Locking, as the reference forces it), a linear
    velocity field v = A x gives every particle -- surface included -- the
    exact strain rate sym(A) and rotation rate skew(A) (to the float weight
    rounding) and the exact density rate -rho tr(A); without it the surface
    particles are far off."""
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    c, d = _jittered(grad_corr=True, nonlock=False, A=A)
    d.calc_dens()
    d.calc_rates()
    scale = np.abs(A).max()
    assert np.abs(d.f("SR") - 0.5 * (A + A.T)).max() < 1e-6 * scale
    assert np.abs(d.f("RR") - 0.5 * (A - A.T)).max() < 1e-6 * scale
    rho = d.f("RHO")
    assert np.abs(d.f("DRHO") + rho * np.trace(A)).max() < 1e-10 * rho.max() * scale
    c, d = _jittered(grad_corr=False, nonlock=False, A=A)
    d.calc_rates()
    assert np.abs(d.f("SR") - 0.5 * (A + A.T)).max() > 0.1 * scale


# ---------------------------------------------------------------- damage
def test_damage_pair_weights_scale_the_locking_passes():
    """Damage enters the Locking acceleration and density passes as the pair
    weight 1 - (D_i + D_j) / 2 (Source/InteractionAlt.cpp:54-60,558-572).
    D = 0 leaves them unchanged, D = 0.5 everywhere halves them exactly (a
    power-of-two weight commutes with rounding), D = 1 removes every pair;
    the rates carry no weight."""
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    out = {}
    for dam in (None, 0.0, 0.5, 1.0):
        c, d = _jittered(nonlock=False, A=A)
        if dam is not None:
            d.set_param("model_damage", 1)
            d.f("DAMD")[:] = dam
        d.calc_accel()
        d.calc_dens()
        d.calc_rates()
        out[dam] = {k: d.f(k).copy() for k in ("A", "DRHO", "SR", "RR")}
    base = out[None]
    assert np.abs(base["A"]).max() > 0 and np.abs(base["DRHO"]).max() > 0
    for k in base:
        assert np.array_equal(out[0.0][k], base[k])
    for k in ("A", "DRHO"):
        assert np.array_equal(out[0.5][k], 0.5 * base[k])
        assert np.all(out[1.0][k] == 0)
    for k in ("SR", "RR"):
        assert np.array_equal(out[0.5][k], base[k]) and np.array_equal(out[1.0][k], base[k])


def _damage_run(D, steps, thermal=False):
    c = cases.damage_small(D, steps)
    d = oracle_from_case(c, threads=4)
    if thermal:
        d.set_param("thermal_solver", 1)
    d.solve_init(c.dt)
    hist = []
    for _ in range(steps):
        d.step()
        S = d.f("SHEAR")[:c.n]
        J2 = 0.5 * (S[:, 0, 0] ** 2 + 2 * S[:, 0, 1] ** 2 + 2 * S[:, 0, 2] ** 2 + S[:, 1, 1] ** 2
                    + 2 * S[:, 1, 2] ** 2 + S[:, 2, 2] ** 2)
        hist.append((d.f("DPLSTRAIN")[:c.n].copy(), np.sqrt(3 * J2), d.f("P")[:c.n].copy(),
                     d.f("PLSTRAIN")[:c.n].copy(), d.f("DAMD")[:c.n].copy()))
    return c, d, hist


def test_damage_accumulates_delta_pl_over_the_fracture_strain():
    """Johnson-Cook damage (Source/Damage.cpp:9-46, Source/Material.cpp:
    76-92) with D2 = D4 = 0: eps_f = D1 (1 + D5 T_h), T = 0 without the
    thermal solver, so T_h = -T_t / (T_m - T_t); every step each particle
    with a positive delta_pl_strain (the last increment, never reset) and
    Sigma_eq > 1e-3 adds delta_pl_strain / eps_f, capped at 1.  The stale
    increments count: a particle keeps gaining damage on steps where its
    plastic strain does not grow.  With damage the neighbour list of the
    first step is never rebuilt, although particles yield."""
    D = (0.05, 0.0, -3.0, 0.0, 0.4)
    c, d, hist = _damage_run(D, 150)
    M = c.materials[0]
    Th = (0.0 - M["jc_Tt"]) / (M["jc_Tm"] - M["jc_Tt"])
    ef = D[0] * (1.0 + D[4] * Th)
    want = np.zeros(c.n)
    stale = 0
    prev_pl = np.zeros(c.n)
    for dpl, seq, _, pl, dam in hist:
        on = (dpl > 0) & (seq > 1e-3)
        want = np.where(on, np.minimum(want + dpl / ef, 1.0), want)
        stale += int(np.sum(on & (pl == prev_pl)))
        prev_pl = pl
        assert np.array_equal(dam, want)
    assert want.max() > 0.01 and stale > 0
    assert np.abs(d.f("PLSTRAIN")).max() > 1e-6 and d.param("nb_searches") == 1


def test_damage_triaxiality_and_rate_terms():
    """The full criterion eps_f = (D1 + D2 exp(D3 s)) (1 + D4 ln(rate/eps0))
    (1 + D5 T_h) with s = -P / Sigma_eq (the triaxiality, the exponential
    only below 1.5) and rate = delta_pl_strain / dt, evaluated from each
    step's fields, and capped at 10."""
    D = (0.05, 0.3, -2.0, 0.02, 0.4)
    c, d, hist = _damage_run(D, 120)
    M = c.materials[0]
    Th = (0.0 - M["jc_Tt"]) / (M["jc_Tm"] - M["jc_Tt"])
    want = np.zeros(c.n)
    for dpl, seq, P, _, dam in hist:
        on = (dpl > 0) & (seq > 1e-3)
        s = np.where(on, -P / np.where(on, seq, 1.0), 0.0)
        rate = dpl / c.dt
        f_sr = np.where(rate > M["jc_eps0"], 1.0 + np.log(np.where(rate > 0, rate, 1.0) / M["jc_eps0"]) * D[3], 1.0)
        f_tr = np.where(s < 1.5, D[0] + D[1] * np.exp(D[2] * s), D[0])
        ef = np.minimum(f_tr * f_sr * (1.0 + D[4] * Th), 10.0)
        want = np.where(on, np.minimum(want + dpl / np.where(on, ef, 1.0), 1.0), want)
        assert np.allclose(dam, want, rtol=1e-13, atol=0)
    assert want.max() > 0.005


# ---------------------------------------------------------------- CalcForce2233
def test_force2233_equals_split_locking_passes():
    """CalcForce2233's acceleration, density rate and rate tensors are the
    Locking passes' terms added pair by pair in the same order: bit for bit
    the split passes' sums (with and without the gradient correction)."""
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    for gc in (False, True):
        c, d = _jittered(nonlock=False, A=A, grad_corr=gc)
        d.calc_accel()
        d.calc_dens()
        d.calc_rates()
        split = {k: d.f(k).copy() for k in ("A", "DRHO", "SR", "RR")}
        d.start_acceleration()
        d.calc_force2233(xsph=0.5, shepard=True)
        for k in split:
            assert np.array_equal(d.f(k), split[k]), (gc, k)


def test_force2233_shepard_xsph_closed_forms():
    """Two particles at distance r: ZWab_1 = (float)(m_2/rho_2) W(r),
    SumDen_1 = m_2 W(r), VXSPH_1 = XSPH m_2 / ((rho_1 + rho_2)/2) W(r)
    (v_2 - v_1); on a jittered billet sum_i m_i VXSPH_i = 0 (each pair's two
    terms cancel) and a uniform velocity gives VXSPH = 0."""
    h, r = 1.2e-3, 1.5e-3
    d = po.Domain(2)
    d.f("X")[:] = [[0.0, 0.0, 0.0], [r, 0.0, 0.0]]
    d.f("H")[:] = h
    d.f("MASS")[:] = [2.0e-6, 3.0e-6]
    d.f("RHO")[:] = [2700.0, 2650.0]
    d.f("REFRHO")[:] = 2700.0
    d.f("V")[:] = [[0.1, -0.2, 0.3], [-0.4, 0.5, 0.2]]
    d.set_material(0, model=po.BILINEAR, E=70e9, nu=0.3)
    d.set_param("nonlock_sum", 0)
    for k in range(3):
        d.set_param(f"DomMax{k}", 0.01)
        d.set_param(f"DomMin{k}", -0.01)
    d.neighbour_search()
    d.start_acceleration()
    xs = 0.7
    d.calc_force2233(xsph=xs, shepard=True)
    W = po.kernel(r / h, h)
    m, rho, v = d.f("MASS"), d.f("RHO"), d.f("V")
    assert d.f("ZWAB")[0] == float(np.float32(m[1] / rho[1])) * W
    assert d.f("ZWAB")[1] == float(np.float32(m[0] / rho[0])) * W
    assert d.f("SUMDEN")[0] == m[1] * W and d.f("SUMDEN")[1] == m[0] * W
    want = xs * m[1] / (0.5 * (rho[0] + rho[1])) * W * (v[1] - v[0])
    assert np.allclose(d.f("VXSPH")[0], want, rtol=1e-15, atol=0)
    A = np.array([[1.0, 2.0, -0.5], [0.3, -1.5, 0.7], [-0.2, 0.4, 0.8]]) * 100.0
    c, d = _jittered(nonlock=False, A=A)
    d.calc_force2233(xsph=0.5, shepard=False)
    mv = (d.f("MASS")[:, None] * d.f("VXSPH")).sum(0)
    assert np.abs(mv).max() < 1e-12 * np.abs(d.f("MASS")[:, None] * d.f("VXSPH")).sum(0).max()
    assert np.all(d.f("SUMDEN") == 0)
    c, d = _jittered(nonlock=False)
    d.f("V")[:] = [0.3, -1.2, 2.0]
    d.calc_force2233(xsph=0.5, shepard=True)
    assert np.all(d.f("VXSPH") == 0)
    # on the (slightly jittered) lattice interior the Shepard sum plus the
    # particle's own term is V sum_k W(|k| dx) over the lattice offsets
    dx, hh = c.meta["dx"], c.h[0]
    rr = np.arange(-3, 4)
    kk = np.stack(np.meshgrid(rr, rr, rr, indexing="ij"), -1).reshape(-1, 3)
    dist = np.sqrt((kk * kk).sum(1)) * dx
    lat = (c.mass[0] / c.density[0]) * sum(po.kernel(q / hh, hh) for q in dist if q < 2 * hh)
    own = d.f("MASS") * np.array([po.kernel(0.0, x) for x in d.f("H")])
    interior = d.i("NB") == d.i("NB").max()
    assert np.allclose((d.f("SUMDEN") + own)[interior] / d.f("RHO")[interior], lat, rtol=1e-2)
