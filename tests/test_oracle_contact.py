"""Pins of the oracle's rigid-mesh contact (Source/Mesh.cpp,
Source/Contact.cpp, Source/Contact_Wang.cpp, Source/Domain.cpp:1415-1491)
against closed forms: mesh geometry, free-surface detection, the penalty
force of a particle under a plane, static and dynamic friction, and the
rigid-body motion of a mesh."""
import math

import numpy as np
import pytest

import pyoracle as po

DX = 1.0e-3
H = 1.2 * DX
RHO = 2700.0
E, NU = 68.9e9, 0.3


def _block(nx=6, ny=6, nz=4, plane_gap=0.4 * H, dens=8, flip=False, v0=(0.0, 0.0, 0.0), mu=(0.0, 0.0)):
    """A cubic-lattice block under a rigid z-plane (normal -z) shifted by a
    quarter cell so no particle projects onto a triangle edge."""
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * DX
    ns = len(x)
    nr = 2 * dens * dens
    d = po.Domain(ns + nr)
    d.set_param("solid_count", ns)
    K = E / (3 * (1 - 2 * NU))
    d.f("X")[:ns] = x
    d.f("H")[:ns] = H
    d.f("MASS")[:ns] = RHO * DX ** 3
    d.f("RHO")[:ns] = d.f("REFRHO")[:ns] = RHO
    d.f("CS")[:ns] = math.sqrt(K / RHO)
    d.f("G")[:ns] = E / (2 * (1 + NU))
    d.f("SIGMAY")[:ns] = 1e12
    d.f("V")[:ns] = np.asarray(v0)
    d.set_material(0, model=po.BILINEAR, E=E, nu=NU)
    ztop = x[:, 2].max()
    z0 = ztop + plane_gap
    # element leg = DX: particles sit a quarter cell off every mesh line
    d.add_plane_mesh(ns, 10, 2, flip, [-0.75 * DX, -0.75 * DX, z0], [(dens - 0.75) * DX, (dens - 0.75) * DX, z0], dens)
    d.set_param("friction_sta", mu[0])
    d.set_param("friction_dyn", mu[1])
    return d, x, ns, z0


def test_plane_mesh_geometry():
    d, x, ns, z0 = _block()
    dl = DX
    n = d.f("NORMAL")[ns:]
    # flipNormals=false: connectivity (n0, n2, n1) gives the -z normal
    assert np.all(n == np.array([0.0, 0.0, -1.0]))
    # right isosceles triangles of leg dl: centroid to the far vertex dl sqrt(5)/3
    r = dl * math.sqrt(5.0) / 3.0
    assert np.allclose(d.f("H")[ns:], float(np.float32(1.1)) * r, rtol=1e-13)
    assert np.allclose(d.f("X")[ns:, 2], z0, rtol=1e-15)
    assert np.all(d.i("ID")[ns:] == 10)
    # centroids: the square's two triangles either side of its n1-n2
    # diagonal have centroids at (1/3, 1/3) and (2/3, 2/3) of the cell
    c = (d.f("X")[ns:, :2] + 0.75 * DX) / dl
    frac = c - np.floor(c + 1e-9)
    assert np.allclose(frac[0::2], 1 / 3, atol=1e-9) and np.allclose(frac[1::2], 2 / 3, atol=1e-9)


def test_free_surface_detection():
    d, x, ns, _ = _block()
    d.solve_init(1e-8)
    ids = d.i("ID")[:ns]
    g = np.round(x / DX - 0.5).astype(int)
    outer = (g == 0).any(1) | (g[:, 0] == 5) | (g[:, 1] == 5) | (g[:, 2] == 3)
    # every particle on the block's faces is free surface; the 2x2x2 interior core
    # (inner with respect to all faces by 2 layers) is not
    assert np.all(ids[outer] == 1)
    core = ((g[:, 0] >= 2) & (g[:, 0] <= 3) & (g[:, 1] >= 2) & (g[:, 1] <= 3) & (g[:, 2] >= 1) & (g[:, 2] <= 2))
    assert np.all(ids[core] == 0)


def test_penalty_force_closed_form():
    """A resting block touching the plane: kij = 2m/dt^2, contact force
    kij (h - dist) n on each top particle, dist = z0 - z_top."""
    dt = 1e-8
    d, x, ns, z0 = _block()
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    kij = 2.0 * m / (dt * dt)
    top = np.isclose(x[:, 2], x[:, 2].max())
    cf = d.f("CONTFORCE")[:ns]
    want = kij * (H - (z0 - x[top, 2]))
    assert np.allclose(cf[top, 2], -want, rtol=1e-12) and np.all(cf[top, :2] == 0)
    assert np.all(cf[~top] == 0)
    total = d.contact_force(0)
    assert total[2] == pytest.approx(cf[:, 2].sum(), rel=1e-12)
    assert d.param("contact_force_sum") == pytest.approx(np.abs(cf[:, 2]).sum(), rel=1e-12)


@pytest.mark.parametrize("mu_s,branch", [(0.2, "static"), (1e-9, "dynamic")])
def test_friction_closed_forms(mu_s, branch):
    """Tangential slip du = v dt of a block sliding along the plane gives
    tgforce = kij v dt.  Static: contforce -= tgforce / m (the reference
    subtracts an acceleration from a force) and a -= tgforce / m.  Dynamic:
    contforce -= mu_d |F_n| t and a -= mu_d |F_n| t / m."""
    dt, vx, mu_d = 1e-8, 0.5, 0.3
    d, x, ns, z0 = _block(v0=(vx, 0.0, 0.0), mu=(mu_s, mu_d))
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    kij = 2.0 * m / (dt * dt)
    top = np.isclose(x[:, 2], x[:, 2].max())
    Fn = kij * (H - (z0 - x[top, 2]))
    tg = kij * vx * dt
    cf = d.f("CONTFORCE")[:ns][top]
    # du = x_pred - x - v_rigid dt cancels x (|x| / |v dt| ~ 1e6): 1e-9 bar
    if branch == "static":
        assert tg < mu_s * Fn.min()
        assert np.allclose(cf[:, 0], -tg / m, rtol=1e-9)
    else:
        assert np.allclose(cf[:, 0], -mu_d * Fn, rtol=1e-12)
    assert np.allclose(cf[:, 2], -Fn, rtol=1e-12)


def test_mesh_motion_and_normal_refresh():
    """A velocity BC on the mesh zone moves nodes and particles by v dt per
    step; the flipped normals of flipNormals=true are recomputed from the
    winding at the first update and so turn over (Source/Mesh.cpp:273-296)."""
    dt, vz = 1e-8, -2.0
    d, x, ns, z0 = _block(plane_gap=3 * H, flip=True)
    d.add_velocity_bc(10, [0.0, 0.0, vz])
    assert np.all(d.f("NORMAL")[ns:] == np.array([0.0, 0.0, -1.0]))
    d.solve_init(dt)
    for _ in range(5):
        d.step()
    assert np.allclose(d.f("X")[ns:, 2], z0 + 5 * vz * dt, rtol=0, atol=1e-15)
    assert np.all(d.f("NORMAL")[ns:] == np.array([0.0, 0.0, 1.0]))
    assert np.allclose(d.f("V")[ns:], [0.0, 0.0, vz], rtol=1e-15)


def test_mesh_rotation_explicit_euler_closed_form():
    """A constant velocity BC with valueAng w = (wx, 0, 0) on the mesh zone:
    TriMesh::Update moves each node by (w x node) dt, an explicit Euler step
    of a rotation, so after k steps every element normal is the initial one
    rotated about x by k atan(wx dt) (the scheme scales the rotated part by
    sqrt(1 + (wx dt)^2) per step, which normalisation removes), and the
    particles take the mean node velocity, w x centroid of the nodes' last
    positions before the update (Source/Mesh.cpp:273-296,
    Source/Contact.cpp:41-55)."""
    dt, wx, k = 1e-6, 2000.0, 7
    d, x, ns, z0 = _block(plane_gap=3 * H)
    d.add_velocity_bc(10, [0.0, 0.0, 0.0])
    d.set_bc_angular(10, [wx, 0.0, 0.0])
    d.solve_init(dt)
    for _ in range(k):
        d.step()
    th = k * math.atan(wx * dt)
    n = d.f("NORMAL")[ns:]
    want = np.array([0.0, math.sin(th), -math.cos(th)])   # R_x(th) (0, 0, -1)
    assert np.abs(n - want).max() < 1e-12
    # rotating about x leaves the x velocity 0 and gives y, z velocities of
    # w x c for the pre-update centroid c: v_y = -wx c_z, v_z = wx c_y
    v = d.f("V")[ns:]
    assert np.all(v[:, 0] == 0.0)
    assert np.abs(v[:, 1]).max() > 0 and np.abs(v[:, 2]).max() > 0


def _run_case(case, steps):
    from oracle_case import oracle_from_case
    d = oracle_from_case(case, threads=4)
    d.solve_init(case.dt)
    for _ in range(steps):
        d.step()
    return d


def test_per_material_friction_and_group_contact_sums():
    """A two-specimen batch (cylinders between their own rigid patches, no
    pair across specimens): with each specimen's friction set per material,
    every specimen's displacements equal bit for bit those of the same batch
    with that specimen's mu as the domain's friction, while the other mu
    changes them (so the friction branch is active and the lookup uses the
    contacting particle's material).  The per-group contact sums add up to
    each mesh's m_contact_force, both groups carry load, and each specimen
    alone gives its group's normal forces (to 1e-6: a lone specimen follows
    its own neighbour-rebuild schedule)."""
    import dataclasses

    from weldform_amd import cases

    steps = 60
    c = cases.config4(n_spec=2, n_range=(600, 900), steps=steps, v_top=-20.0, seed=7)
    mu = c.meta["friction"]
    assert abs(mu[0] - mu[1]) > 0.005
    d = _run_case(c, steps)
    assert np.abs(d.f("PLSTRAIN")).max() > 0
    glob = [_run_case(dataclasses.replace(c, mat_friction={}, contact=(m, m, 0.0)), steps) for m in mu]
    disp = d.f("DISP")[:c.n]
    for i in range(2):
        sel = c.material == i
        assert np.array_equal(disp[sel], glob[i].f("DISP")[:c.n][sel])
        assert np.abs(disp[sel] - glob[1 - i].f("DISP")[:c.n][sel]).max() > 1e-8
    for m in range(2):
        tot = d.contact_force(m)
        gs = [d.group_contact_force(i, m) for i in range(2)]
        assert np.allclose(gs[0] + gs[1], tot, rtol=0, atol=1e-12 * np.abs(tot).max())
        assert all(abs(g[2]) > 0.1 * abs(tot[2]) for g in gs)
    # each specimen alone (its own neighbour-rebuild schedule): the normal
    # force of each patch as its share of the batch
    for i in range(2):
        lone = _run_case(cases.specimen(c, i), steps)
        for m in range(2):
            fz, gz = lone.contact_force(m)[2], d.group_contact_force(i, m)[2]
            assert abs(fz - gz) < 1e-6 * abs(gz)


def _cont_stiff():
    """Fraser 3-153 as SolveDiffUpdateKickDrift forms it: 9 K G / (3 K + G) dS,
    K = Cs^2 rho, dS = (m / rho)^0.33333."""
    K = E / (3 * (1 - 2 * NU))
    Cs = math.sqrt(K / RHO)
    G = E / (2 * (1 + NU))
    bulk = Cs * Cs * RHO
    return 9.0 * bulk * G / (3.0 * bulk + G) * (DX ** 3) ** 0.33333


def _law_block(law, pfac, **kw):
    d, x, ns, z0 = _block(**kw)
    d.set_param("contact_alg", law)
    d.set_param("PFAC", pfac)
    return d, x, ns, z0


@pytest.mark.parametrize("mu_d", [0.0, 0.3])
def test_fraser_law_closed_forms(mu_d):
    """Fraser's law (Source/Contact.cpp:223-554) on a block moving up at vz
    under a resting plane (gap 0.4 h): the top row approaches
    (delta_ = vz), its time to contact t_c = (h - gap) / (-vz) is below dt,
    penetration (dt - t_c) vz = dt vz + h - gap, normal force
    PFAC cont_stiff (dt vz + h - gap); with a tangential velocity above
    1e-3 the dynamic branch ASSIGNS contforce = -mu_d |F| t / m.  The force
    is not applied: the block moves as without the plane."""
    dt, vx, vz, pfac = 1e-8, 0.05, 1.0, 0.7
    d, x, ns, z0 = _law_block(0, pfac, v0=(vx, 0.0, vz), mu=(0.2, mu_d))
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    top = np.isclose(x[:, 2], x[:, 2].max())
    Fn = pfac * _cont_stiff() * (dt * vz + H - (z0 - x[top, 2]))
    cf = d.f("CONTFORCE")[:ns]
    if mu_d == 0.0:
        assert np.allclose(cf[top, 2], -Fn, rtol=1e-9) and np.all(cf[top, :2] == 0)
    else:
        assert np.allclose(cf[top, 0], -mu_d * Fn / m, rtol=1e-9) and np.all(cf[top, 1:] == 0)
    assert np.all(cf[~top] == 0)
    assert d.param("contact_force_sum") == pytest.approx(Fn.sum(), rel=1e-9)
    far, _, _, _ = _law_block(0, pfac, v0=(vx, 0.0, vz), mu=(0.2, mu_d), plane_gap=3 * H)
    far.solve_init(dt)
    far.step()
    assert np.array_equal(d.f("DISP")[:ns], far.f("DISP")[:ns])


def test_fraser_bound_friction_removes_tangential_acceleration():
    """fric_type Fr_Bound: a contacting particle's acceleration loses its
    tangential part (a -= a - (a.n) n), so a sheared block's top row keeps
    its tangential velocity through the first step (v += a dt / 2), while
    without the bound it changes."""
    dt, pfac = 1e-8, 0.7
    g = np.stack(np.meshgrid(np.arange(6), np.arange(6), np.arange(4), indexing="ij"), -1).reshape(-1, 3)
    zc = (g[:, 2] + 0.5) * DX
    v0 = np.stack([200.0 * (zc - zc.mean()), np.zeros_like(zc), np.full_like(zc, 1.0)], axis=1)
    out = {}
    for bound in (0, 1):
        d, x, ns, z0 = _law_block(0, pfac)
        d.f("V")[:ns] = v0
        d.f("ALPHA")[:ns] = 1.0  # Monaghan viscosity on the shear flow: a_t != 0
        d.set_param("fric_bound", bound)
        d.solve_init(dt)
        d.step()
        top = np.isclose(x[:, 2], x[:, 2].max())
        out[bound] = d.f("V")[:ns][top, :2] - v0[top, :2]
    assert np.abs(out[0]).max() > 0
    assert np.all(out[1] == 0)


@pytest.mark.parametrize("vx", [0.0, 0.5])
def test_seo_law_closed_forms(vx):
    """Seo's law (Source/Contact.cpp:562-770): a block under the plane (gap
    0.4 h), normal force PFAC cont_stiff (h - gap) on the top row; with a
    sliding velocity vx and no tangential acceleration the static branch
    holds and contforce -= PFAC cont_stiff vx dt / m.  Not applied."""
    dt, pfac, mu_s = 1e-8, 0.7, 0.2
    d, x, ns, z0 = _law_block(2, pfac, v0=(vx, 0.0, 0.0), mu=(mu_s, mu_s))
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    ks = pfac * _cont_stiff()
    top = np.isclose(x[:, 2], x[:, 2].max())
    Fn = ks * (H - (z0 - x[top, 2]))
    cf = d.f("CONTFORCE")[:ns]
    assert np.allclose(cf[top, 2], -Fn, rtol=1e-12)
    if vx:
        assert np.allclose(cf[top, 0], -ks * vx * dt / m, rtol=1e-9)
    else:
        assert np.all(cf[top, :2] == 0)
    assert np.all(cf[~top] == 0)
    far, _, _, _ = _law_block(2, pfac, v0=(vx, 0.0, 0.0), mu=(mu_s, mu_s), plane_gap=3 * H)
    far.solve_init(dt)
    far.step()
    assert np.array_equal(d.f("DISP")[:ns], far.f("DISP")[:ns])


@pytest.mark.parametrize("vx,mu_s", [(0.0, 0.2), (0.5, 0.2), (0.5, 1e-9)])
def test_lsdyna_law_closed_forms(vx, mu_s):
    """LS-Dyna's law (Source/Contact_LS.cpp:15-292): penalty
    2 m / dt^2 PFAC on the predicted penetration h - gap; friction: the trial
    force -PFAC m v_r / dt, returned to the cone mu_s |F_n|.  Not applied;
    contact_reaction_sum = sum m a.n = 0 here."""
    dt, pfac = 1e-8, 0.7
    d, x, ns, z0 = _law_block(4, pfac, v0=(vx, 0.0, 0.0), mu=(mu_s, mu_s))
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    kij = 2.0 * m / (dt * dt) * pfac
    top = np.isclose(x[:, 2], x[:, 2].max())
    Fn = kij * (H - (z0 - x[top, 2]))
    cf = d.f("CONTFORCE")[:ns]
    trial = pfac * m * vx / dt
    assert np.allclose(cf[top, 2], -Fn, rtol=1e-12)
    if vx == 0.0:
        assert np.all(cf[top, :2] == 0)
    elif trial > mu_s * Fn.min():
        assert np.allclose(cf[top, 0], -mu_s * Fn, rtol=1e-9)
    else:
        assert np.allclose(cf[top, 0], -trial, rtol=1e-12)
    assert np.all(cf[~top] == 0)
    assert d.param("contact_reaction_sum") == 0.0
    assert d.param("contact_force_sum") == pytest.approx(np.linalg.norm(cf[top], axis=1).sum(), rel=1e-12)


def test_linear_friction_function():
    """friction_function Linear: friction_sta = friction_dyn = m T + b of the
    contacting particle (Source/Contact_Wang.cpp:241-244), here in the Wang
    dynamic branch: contforce_t = -mu(T) |F_n|."""
    dt, vx, T0, fm, fb = 1e-8, 0.5, 300.0, 1e-12, 1e-10
    d, x, ns, z0 = _block(v0=(vx, 0.0, 0.0), mu=(5.0, 5.0))
    d.f("T")[:ns] = T0
    d.set_param("friction_linear", 1)
    d.set_param("friction_m", fm)
    d.set_param("friction_b", fb)
    d.solve_init(dt)
    d.step()
    m = RHO * DX ** 3
    kij = 2.0 * m / (dt * dt)
    top = np.isclose(x[:, 2], x[:, 2].max())
    Fn = kij * (H - (z0 - x[top, 2]))
    mu = fm * T0 + fb
    assert kij * vx * dt > mu * Fn.max()  # sliding at mu(T) = 4e-10, static at 5
    cf = d.f("CONTFORCE")[:ns][top]
    assert np.allclose(cf[:, 0], -mu * Fn, rtol=1e-12)
