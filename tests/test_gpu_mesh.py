"""Rigid contact surfaces given as general triangle meshes (wf_add_rigid_mesh,
the TriMesh of a Nastran file, Source/Mesh.cpp:17-64) on the GPU.

  * A plane passed as a general mesh (its nodes, triangles and normals)
    runs bitwise like the same plane built from its parameters
    (wf_add_rigid_plane), FP64 and mixed.
  * A punch tilted by 5 degrees about x (not axis-aligned, so every contact
    test runs off the lattice) against the CPU oracle over the first ~12
    steps of contact: free-surface ids and contact-pair counts exact;
    displacement, plane forces and particle contact forces within
    max(1e-8, 10x the oracle's own sensitivity to a 1e-15 relative change
    of the initial positions) -- the penalty contact is chaotic (see
    test_gpu_contact.py), and the tilted punch more so: that sensitivity
    reaches 1e-8 (displacement) and 9e-8 (contact forces) by step 50."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


def _grid_mesh(p0, e1, e2, dens, flip=False):
    """The nodes and triangles of TriMesh::AxisPlaneMesh's winding over the
    parallelogram p0 + s e1 + t e2 (s, t in [0, 1]); normals from the
    winding, negated when flip (as AxisPlaneMesh's flipNormals)."""
    s = np.arange(dens + 1) / dens
    S, T = np.meshgrid(s, s, indexing="xy")
    nodes = np.asarray(p0)[None, :] + S.reshape(-1, 1) * np.asarray(e1)[None, :] + T.reshape(-1, 1) * np.asarray(e2)
    el = []
    for j in range(dens):
        for i in range(dens):
            n0 = (dens + 1) * j + i
            n1, n2 = n0 + 1, n0 + dens + 1
            n3 = n2 + 1
            el += [(n0, n1, n2), (n1, n3, n2)] if flip else [(n0, n2, n1), (n1, n2, n3)]
    el = np.array(el, dtype=np.int32)
    a, b, c = nodes[el[:, 0]], nodes[el[:, 1]], nodes[el[:, 2]]
    w = np.cross(b - a, c - a)
    nrm = w / np.sqrt((w * w).sum(1))[:, None]
    return nodes, el, (-nrm if flip else nrm)


def _plane_as_mesh(m):
    """A plane dict of cases.contact_small as the same general mesh: the
    nodes as AxisPlaneMesh places them (coordinates accumulated by repeated
    addition of the side dl), its winding and its normals."""
    p1, p2, dens = np.asarray(m["p1"], float), np.asarray(m["p2"], float), m["dens"]
    dl = (p2[0] - p1[0]) / dens
    x, y = np.zeros(dens + 1), np.zeros(dens + 1)
    x[0], y[0] = p1[0], p1[1]
    for i in range(1, dens + 1):
        x[i] = x[i - 1] + dl
        y[i] = y[i - 1] + dl
    X, Y = np.meshgrid(x, y, indexing="xy")
    nodes = np.stack([X.ravel(), Y.ravel(), np.full(X.size, p1[2])], 1)
    _, el, _ = _grid_mesh(p1, (1.0, 0, 0), (0, 1.0, 0), dens, m["flip"])
    a, b, c = nodes[el[:, 0]], nodes[el[:, 1]], nodes[el[:, 2]]
    w = np.cross(b - a, c - a)
    nrm = w / np.sqrt((w * w).sum(1))[:, None]
    if m["flip"]:
        nrm = -nrm
    return dict(zone=m["zone"], nodes=nodes, elems=el, normals=nrm, hfac=m["hfac"])


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_plane_as_general_mesh_is_bitwise_the_plane(precision):
    c = cases.contact_small(steps=120)
    print(f"plane as mesh: n={c.n} precision={precision}", flush=True)
    a = cases.make_solver(c, precision=precision)
    c2 = cases.contact_small(steps=120)
    c2.meshes = [_plane_as_mesh(m) for m in c2.meshes]
    b = cases.make_solver(c2, precision=precision)
    nrig = sum(len(m["elems"]) for m in c2.meshes)
    assert np.array_equal(a.rigid("x", nrig), b.rigid("x", nrig))
    assert np.array_equal(a.rigid("normal", nrig), b.rigid("normal", nrig))
    for g in (a, b):
        g.step(c.steps)
        g.sync()
    assert np.abs(a.contact_force_history(0)).max() > 0
    for f in ("disp", "v", "contforce", "pl_strain"):
        assert np.array_equal(a.get(f), b.get(f)), f
    assert np.array_equal(a.contact_force_history(0), b.contact_force_history(0))
    assert np.array_equal(a.rigid("normal", nrig), b.rigid("normal", nrig))


EARLY = 50


def _tilted_case(steps):
    c = cases.contact_small(steps=steps)
    R, h = c.meta["R"], c.meta["h"]
    zhi = c.x[:, 2].max()
    W = R + 2.5e-3
    t = np.radians(5.0)
    # a square of side 2W in the plane tilted about x, 1.02 h above the
    # top row's particle of least y, which it touches first
    y0 = c.x[c.x[:, 2] > zhi - 0.5 * c.meta["dx"], 1].min()
    p0 = np.array([-W, -W * np.cos(t), zhi + 1.02 * h - (W + y0 / np.cos(t)) * np.sin(t)])
    nodes, el, nrm = _grid_mesh(p0, (2 * W, 0, 0), (0, 2 * W * np.cos(t), 2 * W * np.sin(t)), 22)
    c.meshes = [dict(zone=10, nodes=nodes, elems=el, hfac=1.1), c.meshes[1]]
    assert (nrm[:, 2] < 0).all()   # the punch faces the billet
    c.name = "cyl_upset_4k_contact_tilted_punch"
    return c


def _oracle_run(c, eps=0.0):
    d = oracle_from_case(c, threads=8)
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:c.n] = c.x * (1 + eps * rng.uniform(-1, 1, size=c.x.shape))
    d.solve_init(c.dt)
    F = []
    for _ in range(c.steps):
        d.step()
        F.append(np.stack([d.contact_force(0), d.contact_force(1)]))
    return d, np.array(F)


def test_tilted_mesh_punch_matches_oracle():
    c = _tilted_case(EARLY)
    print(f"tilted punch: n={c.n} steps={c.steps}", flush=True)
    d, F = _oracle_run(c)
    dp, Fp = _oracle_run(c, 1e-15)
    sens = {"disp": normwise_err(dp.f("DISP")[:c.n], d.f("DISP")[:c.n]),
            "F_top": normwise_err(Fp[:, 0], F[:, 0]), "F_bottom": normwise_err(Fp[:, 1], F[:, 1]),
            "contforce": normwise_err(dp.f("CONTFORCE")[:c.n], d.f("CONTFORCE")[:c.n]), "rigid_x": 0.0}
    assert np.abs(F[:, 0, 2]).max() > 0   # the punch touched within the window
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(c.steps)
    g.sync()
    Fg = np.stack([g.contact_force_history(0), g.contact_force_history(1)], axis=1)
    nrig = sum(len(m["elems"]) if "elems" in m else 2 * m["dens"] ** 2 for m in c.meshes)
    errs = {"disp": normwise_err(g.get("disp"), d.f("DISP")[:c.n]),
            "F_top": normwise_err(Fg[:, 0], F[:, 0]),
            "F_bottom": normwise_err(Fg[:, 1], F[:, 1]),
            "contforce": normwise_err(g.get("contforce"), d.f("CONTFORCE")[:c.n]),
            "rigid_x": normwise_err(g.rigid("x", nrig), d.f("X")[c.n:])}
    print("tilted punch errors", errs, "oracle sensitivity", sens, flush=True)
    assert np.array_equal(g.get_int("id"), d.i("ID")[:c.n])
    assert g.scalars()["contact_pairs"] == d.param("ncont")
    for k, e in errs.items():
        assert e < max(1e-8, 10 * sens[k]), k


def test_rotating_punch_matches_oracle():
    """The top plane of the contact case rotating about the x axis
    (valueAng (5, 0, 0) rad/s on its constant velocity BC, TriMesh::Update's
    m_v + m_w x node): element normals and rigid-particle positions and
    velocities after 60 steps within 1e-12 of the oracle; the rest as the
    tilted punch, within max(1e-8, 10x the oracle's own sensitivity)."""
    c = cases.contact_small(steps=60)
    c.bcs = [dict(c.bcs[0], w=(5.0, 0.0, 0.0))]
    c.name = "cyl_upset_4k_contact_rotating_punch"
    print(f"rotating punch: n={c.n} steps={c.steps}", flush=True)
    d, F = _oracle_run(c)
    dp, Fp = _oracle_run(c, 1e-15)
    nrig = sum(2 * m["dens"] ** 2 for m in c.meshes)
    g = cases.make_solver(c, precision=capi.FP64)
    g.step(c.steps)
    g.sync()
    n0 = d.f("NORMAL")[c.n:]
    assert np.abs(n0[:nrig // 2, 1]).max() > 1e-6     # the top plane turned
    for f, key in (("normal", "NORMAL"), ("x", "X"), ("v", "V")):
        e = normwise_err(g.rigid(f, nrig), d.f(key)[c.n:])
        print(f"rotating punch rigid {f}: {e:.3e}", flush=True)
        assert e < 1e-12, f
    Fg = np.stack([g.contact_force_history(0), g.contact_force_history(1)], axis=1)
    errs = {"disp": normwise_err(g.get("disp"), d.f("DISP")[:c.n]), "F_top": normwise_err(Fg[:, 0], F[:, 0]),
            "contforce": normwise_err(g.get("contforce"), d.f("CONTFORCE")[:c.n])}
    sens = {"disp": normwise_err(dp.f("DISP")[:c.n], d.f("DISP")[:c.n]), "F_top": normwise_err(Fp[:, 0], F[:, 0]),
            "contforce": normwise_err(dp.f("CONTFORCE")[:c.n], d.f("CONTFORCE")[:c.n])}
    print("rotating punch errors", errs, "oracle sensitivity", sens, flush=True)
    assert np.abs(F[:, 0, 2]).max() > 0
    assert np.array_equal(g.get_int("id"), d.i("ID")[:c.n])
    for k, e in errs.items():
        assert e < max(1e-8, 10 * sens[k]), k
