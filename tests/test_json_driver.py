"""The WeldForm command-line driver (weldform_amd/driver/weldform.cpp) on
scene files in the reference's JSON format (Source/WeldForm.cpp:147-944):
set-up checks with --dry-run, which builds the domain on the host and stops
before any device work.  The device run of the same file is in
tests/test_gpu_facade.py."""
import copy
import json
import os
import subprocess

import pytest

from weldform_amd import build as wfbuild
from weldform_amd import cases

HERE = os.path.dirname(__file__)
SCENE = os.path.join(HERE, "json", "upsetting_bilinear.json")


def _driver():
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    return wfbuild.JSON_DRIVER


def _dry(tmp_path, scene):
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(scene))
    r = subprocess.run([_driver(), str(p), "--dry-run"], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stdout, r.stderr


def _fields(out):
    words = out.replace(",", " ").split()
    get = lambda k: words[words.index(k) + 1]  # noqa: E731
    return get


def test_scene_builds_config1(tmp_path):
    scene = json.load(open(SCENE))
    rc, out, err = _dry(tmp_path, scene)
    assert rc == 0, err
    c = cases.config1()
    get = _fields(out)
    assert "4524 particles" in out
    assert float(get("dt")) == c.dt                       # cflFactor 0.4: 0.4 h / Cs
    assert out.count("particle count: 156") == 2          # bottom and top rows
    assert float(get("Sigmay0")) == 300e6
    E, Et = 68.9e9, 1.5e9
    assert float(get("Ep")) == E * Et / (E - Et)


def test_material_initial_yield_quirks(tmp_path):
    scene = json.load(open(SCENE))
    # Hollomon: the driver's three-argument CalcYieldStress reaches the base
    # class, so the initial yield stress is 0 (Source/Material.h:98-101,204-224)
    s = copy.deepcopy(scene)
    s["Materials"][0].update(type="Hollomon", const=[7.18e8, 0.22], yieldStress0=260e6)
    rc, out, err = _dry(tmp_path, s)
    assert rc == 0, err
    assert float(_fields(out)("Sigmay0")) == 0.0
    # Johnson-Cook: JC(0, 0, T) with T = 300 from an initial condition of
    # the thermal solver
    s = copy.deepcopy(scene)
    s["Configuration"]["solver"] = "Mech-Thermal-LeapFrog"
    s["Materials"][0].update(type="JohnsonCook", const=[175e6, 380e6, 0.32, 0.0015, 1.0, 1.0, 775.0, 273.0])
    s["InitialConditions"] = [{"Temp": 300.0}]
    rc, out, err = _dry(tmp_path, s)
    assert rc == 0, err
    Th = (300.0 - 273.0) / (775.0 - 273.0)
    assert float(_fields(out)("Sigmay0")) == pytest.approx(175e6 * (1 - Th), rel=1e-15)


@pytest.mark.parametrize("edit,message", [
    (lambda s: s["Configuration"].update(solver="Mech-Fraser"), "Fraser"),
    (lambda s: s["Configuration"].update(solver="Thermal"), "thermal-only"),
    (lambda s: s["DomainBlocks"][0].update(gridCoordSys="Cylindrical"), "Cylindrical"),
    (lambda s: s["DomainBlocks"][0].update(dim=[0.015, 0.0, 0.0005]), "only the 3D generator"),
    (lambda s: s["Materials"][0].update(type="GMT"), "GMT"),
    (lambda s: s["Materials"][0].update(damageModel="Rankine"), "damage"),
    (lambda s: s["Configuration"].update(smoothlenUpdate=True), "smoothlenUpdate"),
    (lambda s: s.update(RigidBodies=[{"type": "Line", "zoneId": 10}]), "Plane and File only"),
    (lambda s: s.update(RigidBodies=[{"type": "File", "zoneId": 10, "fileName": "missing.nas"}]), "cannot open"),
])
def test_scene_features_off_the_path_stop(tmp_path, edit, message):
    s = json.load(open(SCENE))
    edit(s)
    rc, out, err = _dry(tmp_path, s)
    assert rc == 1 and message in err, (out, err)


def test_thin_box_scene_is_plane_strain(tmp_path):
    """A Box block thinner than h is 2D (Source/WeldForm.cpp:420-424): the
    square packing of pitch 2r in the plane, 14 x 14 particles for a 15 mm
    square at r = 0.5 mm (the generator stops a centre short of L - r)."""
    s = json.load(open(SCENE))
    s["DomainBlocks"][0].update(type="Box", dim=[0.015, 0.015, 0.0])
    rc, out, err = _dry(tmp_path, s)
    assert rc == 0, err
    assert "Dimension is set to 2" in out and "196 particles (196 solid)" in out


def test_contact_scene_adds_rigid_particles(tmp_path):
    s = json.load(open(SCENE))
    s["RigidBodies"] = [{"type": "Plane", "zoneId": 10, "start": [-0.01, -0.01, 0.0296], "dim": [0.02, 0.02, 0.0],
                         "partSide": 20}]
    s["Contact"] = [{"fricCoeffStatic": 0.2, "fricCoeffDynamic": 0.2, "penaltyFactor": 0.6}]
    s["BoundaryConditions"].append({"zoneId": 10, "valueType": 0, "value": [0.0, 0.0, -1.0]})
    rc, out, err = _dry(tmp_path, s)
    assert rc == 0, err
    assert f"{4524 + 2 * 20 * 20} particles (4524 solid)" in out and "contact 1" in out and "meshes 1" in out
    assert "contact heat conduction 0 (hc 0), friction 0" in out
    # contact heat (Source/WeldForm.cpp:711-723) and the punch's temperature
    # from an initial condition on its zone (:850-866)
    s["Configuration"]["solver"] = "Mech-Thermal-LeapFrog"
    s["Contact"][0].update(heatConductance=True, heatCondCoeff=5.0e4, heatFriction=True)
    s["InitialConditions"] = [{"Temp": 300.0}, {"Temp": 450.0, "zoneId": 10}]
    s["Configuration"]["autoTS"] = [True, False, True]
    rc, out, err = _dry(tmp_path, s)
    assert rc == 0, err
    assert "contact heat conduction 1 (hc 50000), friction 1, rigid T 450 .. 450" in out
    assert "autoTS [1 0 1]" in out
