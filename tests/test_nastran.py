"""Nastran triangle meshes (the reference's NastranReader,
Source/NastranReader.h:26-204, and TriMesh(NastranReader &, flip),
Source/Mesh.cpp:17-64) on the CPU, through the facade headers: fixed
8-character fields, the short exponent forms ("0.5633-18", "-.962-03",
"1.5+2"), the leading-'+' quirk (the inserted E leaves "E+1.5", which reads
as 0), node ids mapped to positions, flipNormals swapping the first two
nodes while the first normals keep the file's winding; and a scene with a
"File" rigid body in the command-line driver."""
import json
import os
import subprocess

import numpy as np

from weldform_amd import build as wfbuild

HERE = os.path.dirname(os.path.abspath(__file__))


def _card(*fields):
    return "".join(f"{f:<8}" if i == 0 else f"{f:>8}" for i, f in enumerate(fields))


# (text of the x field, value the reader gives)
XS = [(".5633-18", 0.5633e-18), ("-.962-03", -0.962e-3), ("-.004600", -0.0046), ("1.5+2", 150.0),
      ("+1.5", 0.0), (".0013564", 0.0013564), ("12.", 12.0)]


def _write_mesh(path, ids=(11, 12, 13, 14), z=0.0):
    """Four nodes (ids not starting at 1), two triangles (ids in the
    element cards)."""
    lines = ["$ test mesh", "BEGIN BULK"]
    P = [("0.0", "0.0", f"{z}"), ("1.0", "0.0", f"{z}"), ("0.0", "1.0", f"{z}"), ("1.0", "1.0", f"{z}")]
    for nid, (x, y, zz) in zip(ids, P):
        lines.append(_card("GRID", nid, 0, x, y, zz, 0))
    lines.append(_card("CTRIA3", 1, 1, ids[0], ids[1], ids[2]))
    lines.append(_card("CTRIA3", 2, 1, ids[1], ids[3], ids[2]))
    lines.append("ENDDATA")
    path.write_text("\n".join(lines) + "\n")


def _probe(tmp_path):
    exe = str(tmp_path / "nastran_probe")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(wfbuild.ROOT, "include"),
                    os.path.join(HERE, "cpp", "nastran_probe.cpp"), "-o", exe], check=True)
    return exe


def _read(exe, path, flip):
    r = subprocess.run([exe, str(path), str(int(flip))], capture_output=True, text=True, timeout=30)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    nn, ne = map(int, lines[0].split())
    nodes = np.array([[float(v) for v in lines[1 + k].split()] for k in range(nn)])
    el = np.array([[float(v) for v in lines[1 + nn + e].split()] for e in range(ne)])
    return nodes, el[:, :3].astype(int), el[:, 3:6], el[:, 6:9]


def test_short_exponent_fields(tmp_path):
    exe = _probe(tmp_path)
    xs = XS
    lines = []
    for k, (txt, _) in enumerate(xs):
        lines.append(_card("GRID", k + 1, 0, txt, "0.0", "0.0"))
    lines.append(_card("CTRIA3", 1, 1, 1, 2, 3))
    p = tmp_path / "fields.nas"
    p.write_text("\n".join(lines) + "\n")
    nodes, _, _, _ = _read(exe, p, False)
    assert nodes[:, 0].tolist() == [v for _, v in xs]


def test_mesh_ids_winding_and_flip(tmp_path):
    exe = _probe(tmp_path)
    p = tmp_path / "square.nas"
    _write_mesh(p)
    nodes, el, cen, nrm = _read(exe, p, False)
    assert nodes.tolist() == [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]]
    assert el.tolist() == [[0, 1, 2], [1, 3, 2]]
    assert np.allclose(cen, [[1 / 3, 1 / 3, 0], [2 / 3, 2 / 3, 0]], rtol=0, atol=1e-16)
    assert nrm.tolist() == [[0, 0, 1], [0, 0, 1]]
    # flipNormals swaps the first two nodes; the first normals keep the
    # file's winding (TriMesh(NastranReader &, flip) takes them from elcon)
    _, elf, cenf, nrmf = _read(exe, p, True)
    assert elf.tolist() == [[1, 0, 2], [3, 1, 2]]
    assert np.array_equal(cenf, cen) and nrmf.tolist() == [[0, 0, 1], [0, 0, 1]]


def test_unknown_node_and_missing_file_fail(tmp_path):
    exe = _probe(tmp_path)
    p = tmp_path / "bad.nas"
    p.write_text(_card("GRID", 1, 0, "0.0", "0.0", "0.0") + "\n" + _card("CTRIA3", 1, 1, 1, 2, 3) + "\n")
    r = subprocess.run([exe, str(p), "0"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "unknown node" in r.stderr
    r = subprocess.run([exe, str(tmp_path / "none.nas"), "0"], capture_output=True, text=True, timeout=30)
    assert r.returncode == 1 and "cannot open" in r.stderr


def test_scene_with_file_rigid_body(tmp_path):
    """A punch read from a Nastran file, scaled and translated, in the
    reference's scene format: the dry run adds one rigid particle per
    triangle."""
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    mesh = tmp_path / "punch.nas"
    # 3 x 3 squares of 10 mm, two triangles each, in the z = 0 plane
    lines = []
    nid = lambda i, j: 100 + 4 * j + i  # noqa: E731
    for j in range(4):
        for i in range(4):
            lines.append(_card("GRID", nid(i, j), 0, f"{-15.0 + 10 * i:.1f}", f"{-15.0 + 10 * j:.1f}", "0.0"))
    e = 1
    for j in range(3):
        for i in range(3):
            lines.append(_card("CTRIA3", e, 1, nid(i, j), nid(i, j + 1), nid(i + 1, j)))
            lines.append(_card("CTRIA3", e + 1, 1, nid(i + 1, j), nid(i, j + 1), nid(i + 1, j + 1)))
            e += 2
    mesh.write_text("\n".join(lines) + "\n")
    scene = json.load(open(os.path.join(HERE, "json", "upsetting_bilinear.json")))
    scene["RigidBodies"] = [{"type": "File", "fileName": str(mesh), "zoneId": 10, "scaleFactor": 1.0e-3,
                             "translation": [0.0, 0.0, 0.0312], "flipNormals": False}]
    scene["Contact"] = [{"fricCoeffStatic": 0.2, "fricCoeffDynamic": 0.2, "penaltyFactor": 0.6}]
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(scene))
    r = subprocess.run([wfbuild.JSON_DRIVER, str(p), "--dry-run"], capture_output=True, text=True, timeout=60,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Reading Mesh input file" in r.stdout
    head = r.stdout.split("WeldForm (device path): ")[1]
    total, solid = int(head.split()[0]), int(head.split("particles (")[1].split()[0])
    assert total - solid == 18
