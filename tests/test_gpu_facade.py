"""The C++ SPH::Domain facade (include/weldform) against the CPU oracle.

tests/cpp/upsetting.cpp builds the 4,524-particle bilinear upsetting case the
way a reference driver does (AddCylinderLength, per-particle material
fields, AssignZone, bConds, SolveDiffUpdateLeapFrog / KickDrift), runs it on
the device and writes the zone-3 reaction history and the final particle
state.  The same case goes through the oracle here; the bar is the FP64 one
of the direct C-ABI tests (1e-8 after the run).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import h5read
from oracle_case import normwise_err, oracle_from_case
from weldform_amd import build as wfbuild
from weldform_amd import capi, cases, output

pytestmark = pytest.mark.gpu


def _driver():
    if not os.path.exists(wfbuild.DRIVER):
        wfbuild.build_driver()
    return wfbuild.DRIVER


def _run_driver(tmp_path, steps, scheme, precision, ts="fixed"):
    key = str(tmp_path / "up")
    print(f"facade driver: {steps} steps {scheme} {precision} {ts}", flush=True)
    r = subprocess.run([_driver(), str(steps), key, scheme, precision, ts], capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    react = np.loadtxt(key + "_reaction.txt").reshape(-1, 3)
    final = np.loadtxt(key + "_final.txt")
    if ts == "auto":
        words = r.stdout.split()
        return key, react, final, float(words[words.index("time") + 1])
    return key, react, final


def _oracle(steps, scheme, grad_corr=False):
    case = cases.config1(steps=steps)
    d = oracle_from_case(case, threads=8)
    if grad_corr:
        # the reference's solver switches to the Locking sums with the correction
        d.set_param("nonlock_sum", 0)
        d.set_param("grad_corr", 1)
    d.solve_init(case.dt)
    R = []
    for _ in range(steps):
        d.step(scheme)
        R.append(d.reaction(case.top_zone).copy())
    return case, d, np.array(R)


@pytest.mark.parametrize("scheme,steps", [("leapfrog", 400), ("kickdrift", 200)])
def test_facade_driver_matches_oracle(tmp_path, scheme, steps):
    key, react, final = _run_driver(tmp_path, steps, scheme, "fp64")
    case, d, R = _oracle(steps, scheme)
    assert react.shape == (steps, 3)
    assert final.shape == (case.n, 5)
    e_r = normwise_err(react[:, 2], R[:, 2])
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    e_p = normwise_err(final[:, 4], d.f("PLSTRAIN")) if np.abs(d.f("PLSTRAIN")).max() > 0 else 0.0
    print(f"facade {scheme}: reaction {e_r:.3e} displacement {e_u:.3e} pl_strain {e_p:.3e}", flush=True)
    assert e_r < 1e-8 and e_u < 1e-8 and e_p < 1e-8
    # output files of the solve loop: one after step 1, one at steps/2
    for idx in (1, 2):
        fn = f"{key}_{idx:04d}.csv"
        assert os.path.exists(fn), fn
        with open(fn) as f:
            head = f.readline()
            rows = sum(1 for _ in f)
        assert head.startswith("X, Y, Z, ID, Sigma_eq, Pl_Strain")
        assert rows == case.n


@pytest.mark.parametrize("scheme,steps", [("leapfrog", 300), ("kickdrift", 150)])
def test_host_hook_driver_matches_oracle(tmp_path, scheme, steps):
    """A driver whose BCs are a host GeneralAfter (Compression.cpp's form,
    tests/cpp/compression_hooks.cpp: a ramped top-row velocity, a fixed
    bottom row, its own reaction observer) runs on the facade's slow path,
    the step cut at the reference's GeneralBefore / GeneralAfter call sites.
    The oracle runs the same BCs as UserAcc zones with the equivalent
    amplitude; FP64 bar 1e-8 after the run, as the other facade tests."""
    if not os.path.exists(wfbuild.HOOK_DRIVER):
        wfbuild.build_hook_driver()
    key = str(tmp_path / "hooks")
    print(f"hook driver: {steps} steps {scheme}", flush=True)
    r = subprocess.run([wfbuild.HOOK_DRIVER, str(steps), key, scheme], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    assert int(words[words.index("GeneralBefore") - 1]) == steps
    assert int(words[words.index("GeneralAfter") - 1]) == 3 * steps
    react = np.loadtxt(key + "_reaction.txt").reshape(-1, 3)
    final = np.loadtxt(key + "_final.txt")
    case = cases.config1(steps=steps)
    tau = 50 * case.dt
    case.amps = [dict(id=1, t=[0.0, tau, 1.0], v=[0.0, 1.0, 1.0])]
    case.bcs = [dict(zone=2, value=(0.0, 0.0, 0.0)), dict(zone=3, value=(0.0, 0.0, -1.0), value_type=1, amp_id=1)]
    d = oracle_from_case(case, threads=8)
    d.solve_init(case.dt)
    R = []
    for _ in range(steps):
        d.step(scheme)
        R.append(d.reaction(case.top_zone).copy())
    R = np.array(R)
    assert react.shape == R.shape
    e_r = normwise_err(react[:, 2], R[:, 2])
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    e_p = normwise_err(final[:, 4], d.f("PLSTRAIN")) if np.abs(d.f("PLSTRAIN")).max() > 0 else 0.0
    print(f"hook driver {scheme}: reaction {e_r:.3e} displacement {e_u:.3e} pl_strain {e_p:.3e}", flush=True)
    assert e_r < 1e-8 and e_u < 1e-8 and e_p < 1e-8


def test_facade_driver_mixed_precision(tmp_path):
    steps = 300
    _, react, final = _run_driver(tmp_path, steps, "leapfrog", "mixed")
    _, d, R = _oracle(steps, "leapfrog")
    e_r = normwise_err(react[:, 2], R[:, 2])
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    print(f"facade mixed: reaction {e_r:.3e} displacement {e_u:.3e}", flush=True)
    assert e_r < 1e-4 and e_u < 1e-4


def test_facade_driver_adaptive_step(tmp_path):
    """Domain::auto_ts through the facade: the solve loop runs on the device
    clock (Time <= tf with the adapted steps), outputs follow Time >= tout."""
    steps = 300
    key, react, final, t_end = _run_driver(tmp_path, steps, "leapfrog", "fp64", "auto")
    case = cases.config1(steps=steps)
    tf = (steps - 0.5) * case.dt
    case.auto_ts = 1
    dt0 = case.dt / 0.4   # h / Cs
    d = oracle_from_case(case, threads=8)
    d.solve_init(dt0)
    R = []
    while d.param("Time") <= tf:
        d.step("leapfrog")
        R.append(d.reaction(case.top_zone).copy())
    R = np.array(R)
    print(f"facade adaptive: {len(react)} steps (oracle {len(R)}), time {t_end!r} ({d.param('Time')!r})",
          flush=True)
    # the step settles at the velocity limit 0.7 dx / Cs = 0.58 h / Cs, above
    # the fixed 0.4 h / Cs, so fewer steps cover the same time
    assert len(R) < 0.8 * steps
    assert react.shape == R.shape
    assert t_end == pytest.approx(d.param("Time"), rel=1e-12)
    e_r = normwise_err(react[:, 2], R[:, 2])
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    print(f"facade adaptive: reaction {e_r:.3e} displacement {e_u:.3e}", flush=True)
    assert e_r < 1e-8 and e_u < 1e-8
    assert os.path.exists(f"{key}_0001.csv") and os.path.exists(f"{key}_0002.csv")


@pytest.mark.parametrize("grad_corr", [False, True])
def test_json_scene_driver_matches_oracle(tmp_path, grad_corr):
    """The WeldForm command-line driver on tests/json/upsetting_bilinear.json
    (config 1 in the reference's scene format: 200 Leapfrog steps) against
    the oracle's config-1 run; also with "kernelGradCorr": true, as the
    reference's Compression.json sets it."""
    scene = os.path.join(os.path.dirname(__file__), "json", "upsetting_bilinear.json")
    if grad_corr:
        sc = json.load(open(scene))
        sc["Configuration"]["kernelGradCorr"] = True
        scene = str(tmp_path / "upsetting_gradcorr.json")
        with open(scene, "w") as f:
            json.dump(sc, f)
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    dump = str(tmp_path / "final.txt")
    print(f"json driver: upsetting_bilinear.json kernelGradCorr={grad_corr}", flush=True)
    r = subprocess.run([wfbuild.JSON_DRIVER, scene, "--out", str(tmp_path / "run"), "--dump", dump],
                       capture_output=True, text=True, timeout=100, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    if grad_corr:
        assert "Summation changed to Locking" in r.stdout
    final = np.loadtxt(dump)
    case, d, R = _oracle(200, "leapfrog", grad_corr=grad_corr)
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    print(f"json driver: displacement {e_u:.3e}", flush=True)
    assert final.shape == (case.n, 4) and e_u < 1e-8
    assert os.path.exists(str(tmp_path / "run_0001.csv"))
    if output.has_hdf5():
        # the solve's XDMF files (initial model, each output) and the final
        # state from the Particle objects, maz.hdf5 (Source/WeldForm.cpp:941)
        for k in ("run_Initial", "run_0001", "maz"):
            assert os.path.exists(str(tmp_path / (k + ".xmf"))), k
        ds = h5read.read(str(tmp_path / "maz.hdf5"))
        assert ds["NP"][1][0] == case.n
        assert np.array_equal(ds["Displacement"][1].reshape(-1, 3), final[:, :3].astype(np.float32))


def test_json_contact_heat_scene_runs(tmp_path):
    """The driver on config 1's scene with a punch plane at 450 K (an
    initial condition on its zone), contact conductance and friction
    heating: the facade hands the rigid temperatures and heat settings to
    the device, so heat flows into the 300 K billet."""
    s = json.load(open(os.path.join(os.path.dirname(__file__), "json", "upsetting_bilinear.json")))
    s["Configuration"]["solver"] = "Mech-Thermal-LeapFrog"
    s["Materials"][0].update(thermalCond=237.0, thermalHeatCap=897.0)
    s["RigidBodies"] = [{"type": "Plane", "zoneId": 10, "start": [-0.01, -0.01, 0.0296], "dim": [0.02, 0.02, 0.0],
                         "partSide": 20}]
    s["Contact"] = [{"fricCoeffStatic": 0.2, "fricCoeffDynamic": 0.2, "penaltyFactor": 0.6,
                     "heatConductance": True, "heatCondCoeff": 5.0e4, "heatFriction": True}]
    s["BoundaryConditions"].append({"zoneId": 10, "valueType": 0, "value": [0.0, 0.0, -1.0]})
    s["InitialConditions"] = [{"Temp": 300.0}, {"Temp": 450.0, "zoneId": 10}]
    scene = str(tmp_path / "contact_heat.json")
    with open(scene, "w") as f:
        json.dump(s, f)
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    print("json driver: contact heat scene", flush=True)
    r = subprocess.run([wfbuild.JSON_DRIVER, scene], capture_output=True, text=True, timeout=100, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "rigid T 450 .. 450" in r.stdout
    flux = float(r.stdout.split("Total contact heat flux")[1].split()[0])
    fric = float(r.stdout.split("Contact friction work")[1].split()[0])
    print(f"json driver: contact heat {flux:.6g} J, friction heat {fric:.6g} J", flush=True)
    assert flux > 0 and np.isfinite(flux) and fric >= 0 and np.isfinite(fric)


def test_json_damage_scene_matches_oracle(tmp_path):
    """The driver on config 1's scene with a Johnson-Cook material and
    "damageModel": "JohnsonCook" (the reference's Compression_damage.json
    form; T_t = 0 since the mechanical scene leaves T at 0), zones three rows
    deep, top driven at -20 m/s, 200 Leapfrog steps: the driver switches to
    the Locking sums (Source/WeldForm.cpp:303,314) and runs the damage on
    the device; displacement and plastic strain against the oracle's run of
    the same case."""
    steps = 200
    sc = json.load(open(os.path.join(os.path.dirname(__file__), "json", "upsetting_bilinear.json")))
    jc = [175e6, 380e6, 0.32, 0.0015, 1.0, 1.0, 775.0, 0.0]
    D = [0.2, 0.3, -2.0, 0.01, 0.4]
    sc["Materials"][0].update({"type": "JohnsonCook", "const": jc, "damageModel": "JohnsonCook",
                               "damageParams": D})
    c = cases.config1(steps=steps)
    dx = c.meta["dx"]
    z = c.x[:, 2]
    sc["DomainZones"] = [{"id": 2, "start": [-1.0, -1.0, -1.0], "end": [1.0, 1.0, z.min() + 2.5 * dx]},
                         {"id": 3, "start": [-1.0, -1.0, z.max() - 2.5 * dx], "end": [1.0, 1.0, 1.0]}]
    sc["BoundaryConditions"][1]["value"] = [0.0, 0.0, -20.0]
    sc["Configuration"]["simTime"] = (steps - 0.5) * c.dt
    sc["Configuration"]["outTime"] = 1.0
    scene = str(tmp_path / "upsetting_damage.json")
    with open(scene, "w") as f:
        json.dump(sc, f)
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    dump = str(tmp_path / "final.txt")
    print("json driver: JC damage scene", flush=True)
    r = subprocess.run([wfbuild.JSON_DRIVER, scene, "--out", str(tmp_path / "run"), "--dump", dump],
                       capture_output=True, text=True, timeout=100, cwd=str(tmp_path))
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Assigned Johnson Cook Damage Model" in r.stdout
    final = np.loadtxt(dump)
    mat = dict(cases.JC_AL, jc_Tt=0.0, E=c.materials[0]["E"], nu=c.materials[0]["nu"], G=c.materials[0]["G"],
               Cs=c.materials[0]["Cs"], ref_density=c.materials[0]["ref_density"])
    c.materials = {0: mat}
    c.sigmay = np.full(c.n, jc[0])
    c.ids = np.where(z < z.min() + 2.5 * dx, 2, np.where(z > z.max() - 2.5 * dx, 3, 0)).astype(np.int32)
    c.bcs = [dict(zone=2, value=(0.0, 0.0, 0.0)), dict(zone=3, value=(0.0, 0.0, -20.0))]
    c.damage = {0: tuple(D)}
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    for _ in range(steps):
        d.step()
    assert d.f("DAMD").max() > 0.01
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    e_p = normwise_err(final[:, 3], d.f("PLSTRAIN"))
    print(f"json damage driver: displacement {e_u:.3e} plastic strain {e_p:.3e}", flush=True)
    # the oracle's own sensitivity to a 1e-15 perturbation of the initial
    # positions: 7e-11 displacement, 3e-9 plastic strain
    assert final.shape == (c.n, 4) and e_u < 1e-8 and e_p < 1e-7


def test_json_plane_strain_box_scene_matches_oracle(tmp_path):
    """A Box block thinner than h makes the scene plane strain (Source/
    WeldForm.cpp:420-424): the driver's 2D AddBoxLength square lattice
    (pitch 2r, mass 4 r^2 rho, Source/Domain.cpp:550-571) upset along y for
    200 Leapfrog steps, against the oracle's 2D run of the same lattice."""
    steps = 200
    sc = json.load(open(os.path.join(os.path.dirname(__file__), "json", "upsetting_bilinear.json")))
    r, rho = 0.0005, 2700.0
    Lx, Ly = 0.015, 0.030
    sc["DomainBlocks"] = [{"type": "Box", "zoneId": 0, "start": [0.0, 0.0, 0.0], "dim": [Lx, Ly, 0.0]}]
    sc["DomainZones"] = [{"id": 2, "start": [-1.0, -1.0, -1.0], "end": [1.0, 0.00075, 1.0]},
                         {"id": 3, "start": [-1.0, Ly - 0.00175, -1.0], "end": [1.0, 1.0, 1.0]}]
    sc["BoundaryConditions"] = [{"zoneId": 2, "valueType": 0, "value": [0.0, 0.0, 0.0]},
                                {"zoneId": 3, "valueType": 0, "value": [0.0, -1.0, 0.0]}]
    E, nu, Et = 68.9e9, 0.3, 1.5e9
    K, G, Cs = cases.elastic(E, nu, rho)
    h = 1.2 * 2 * r
    dt = 0.4 * h / Cs
    sc["Configuration"]["simTime"] = (steps - 0.5) * dt
    sc["Configuration"]["outTime"] = 1.0
    scene = str(tmp_path / "plate2d.json")
    with open(scene, "w") as f:
        json.dump(sc, f)
    if not os.path.exists(wfbuild.JSON_DRIVER):
        wfbuild.build_json_driver()
    dump = str(tmp_path / "final.txt")
    print("json driver: plane-strain box scene", flush=True)
    run = subprocess.run([wfbuild.JSON_DRIVER, scene, "--out", str(tmp_path / "run"), "--dump", dump],
                         capture_output=True, text=True, timeout=100, cwd=str(tmp_path))
    assert run.returncode == 0, run.stdout + run.stderr
    assert "Dimension is set to 2" in run.stdout
    final = np.loadtxt(dump)
    # the same lattice, in the generator's order (x fastest) and with its
    # loop bounds evaluated as it evaluates them (xp <= Lx - r in doubles)
    def count(L):
        k = 0
        while (0.0 if k == 0 else 0.0 + (2.0 * k + 1) * r) <= 0.0 + L - r:
            k += 1
        return k
    nx, ny = count(Lx), count(Ly)
    gy, gx = np.meshgrid((2 * np.arange(ny) + 1) * r, (2 * np.arange(nx) + 1) * r, indexing="ij")
    x = np.stack([gx.ravel(), gy.ravel(), np.zeros(nx * ny)], axis=1)
    n = len(x)
    assert final.shape[0] == n
    ids = np.where(x[:, 1] <= 0.00075, 2, np.where(x[:, 1] >= Ly - 0.00175, 3, 0)).astype(np.int32)
    mat = dict(model=capi.BILINEAR, alpha=1.0, beta=0.0, Ep=E * Et / (E - Et), E=E, nu=nu, G=G, Cs=Cs,
               ref_density=rho, P0=0.0)
    c = cases.Case(name="plate2d_json", x=x, h=np.full(n, h), mass=np.full(n, 4 * r * r * rho),
                   density=np.full(n, rho), ids=ids, material=np.zeros(n, dtype=np.int32), materials={0: mat},
                   sigmay=np.full(n, 300e6), dt=dt, steps=steps,
                   bcs=[dict(zone=2, value=(0.0, 0.0, 0.0)), dict(zone=3, value=(0.0, -1.0, 0.0))], dim=2,
                   dom_max=(0.0, -1e11, -1e11), dom_min=(0.0, 1e11, 1e11))
    d = oracle_from_case(c, threads=8)
    d.solve_init(c.dt)
    for _ in range(steps):
        d.step()
    e_u = normwise_err(final[:, :3], d.f("DISP"))
    print(f"json plane-strain driver: displacement {e_u:.3e}", flush=True)
    assert e_u < 1e-8
    assert np.all(final[:, 2] == 0.0)
