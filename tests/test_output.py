"""Output writers on the CPU (libweldform_io.so and the host packing):
Domain::WriteXDMF / Domain::WriteCSV (Source/Output.cpp:75-554).

  * Sigma_eq (Particle::CalculateEquivalentStress) against the von Mises
    closed form sqrt(3/2 s:s) and the uniaxial case sigma_eq = |sigma|;
  * the .hdf5 read back through the HDF5 library: every dataset of the
    reference (names, 1-D sizes n*width, float32 / int32, /NP = n) holding
    exactly the packed values;
  * the .xmf parsed as XML: a Polyvertex grid of n points, every DataItem
    pointing at a dataset of the .hdf5 whose size matches its Dimensions;
  * the .csv byte for byte equal to std::ostream's default formatting of
    the same doubles (a C++ program built here writes it), and the same for
    every thread count."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import h5read
from weldform_amd import build, capi, output

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sigma_eq_pins():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((200, 3, 3)) * 1e8
    S = 0.5 * (A + A.transpose(0, 2, 1))
    S -= np.trace(S, axis1=1, axis2=2)[:, None, None] * np.eye(3) / 3.0
    vm = np.sqrt(1.5 * np.einsum("nij,nij->n", S, S))
    assert np.allclose(output.sigma_eq(S), vm, rtol=1e-14, atol=0)
    # uniaxial stress sigma: deviator diag(2/3, -1/3, -1/3) sigma -> |sigma|
    sig = rng.uniform(-5e8, 5e8, 50)
    U = np.zeros((50, 3, 3))
    U[:, 0, 0], U[:, 1, 1], U[:, 2, 2] = 2 * sig / 3, -sig / 3, -sig / 3
    assert np.allclose(output.sigma_eq(U), np.abs(sig), rtol=1e-14)
    # pure shear tau: sqrt(3) |tau|
    T = np.zeros((50, 3, 3))
    T[:, 0, 1] = T[:, 1, 0] = sig
    assert np.allclose(output.sigma_eq(T), np.sqrt(3) * np.abs(sig), rtol=1e-14)


def _random_fields(n, seed=1):
    rng = np.random.default_rng(seed)
    t = lambda: rng.standard_normal((n, 3, 3))  # noqa: E731
    sym = lambda m: 0.5 * (m + m.transpose(0, 2, 1))  # noqa: E731
    return {"x": rng.standard_normal((n, 3)), "v": rng.standard_normal((n, 3)), "a": rng.standard_normal((n, 3)),
            "disp": rng.standard_normal((n, 3)), "pressure": rng.standard_normal(n),
            "density": rng.uniform(7000, 8000, n), "mass": rng.uniform(1e-6, 2e-6, n), "h": rng.uniform(1e-3, 2e-3, n),
            "sigma": sym(t()) * 1e8, "shear": sym(t()) * 1e8, "strain": sym(t()), "strain_rate": sym(t()),
            "strain_pl": sym(t()), "T": rng.uniform(290, 900, n), "pl_strain": rng.uniform(0, 1, n),
            "ps_energy": rng.uniform(0, 1, n), "contforce": rng.standard_normal((n, 3)),
            "id": rng.integers(0, 4, n), "nb": rng.integers(0, 60, n), "contnb": rng.integers(0, 4, n)}


@pytest.mark.skipif(not h5read.available(), reason="no HDF5 library in this image")
def test_xdmf_files(tmp_path):
    assert output.has_hdf5()
    n = 257
    fl = _random_fields(n)
    f, iv, _ = output.table_from_fields(fl)
    key = str(tmp_path / "dom_0001")
    output.write_xdmf_table(key, n, f, iv, prop_names=("Sy", "Property2", "Damage_like"))
    ds = h5read.read(key + ".hdf5")
    want = output.datasets(f, iv, n)
    want["Property1"], want["Property3"] = want.pop("Property1"), want.pop("Property3")
    names = set(want) - {"Property1", "Property3"} | {"Sy", "Damage_like", "NP"}
    assert set(ds) == names
    assert ds["NP"][0] == (1,) and ds["NP"][1][0] == n
    for name, arr in want.items():
        key_name = {"Property1": "Sy", "Property3": "Damage_like"}.get(name, name)
        shape, got = ds[key_name]
        assert shape == (arr.size,), name
        assert got.dtype == arr.dtype and np.array_equal(got, arr), name
    # the values themselves: positions, the Sigma-as-ShearS quirk, tensor order
    assert np.array_equal(ds["Position"][1].reshape(n, 3), fl["x"].astype(np.float32))
    sig6 = ds["Sigma"][1].reshape(n, 6)
    assert np.array_equal(ds["ShearS"][1], ds["Sigma"][1])
    assert np.array_equal(sig6[:, 4], fl["sigma"][:, 1, 2].astype(np.float32))
    assert np.array_equal(sig6[:, 5], fl["sigma"][:, 0, 2].astype(np.float32))
    assert np.array_equal(ds["Tag"][1], fl["id"].astype(np.int32))
    # the .xmf
    root = ET.parse(key + ".xmf").getroot()
    assert root.tag == "Xdmf" and root.get("Version") == "2.0"
    grid = root.find("Domain/Grid")
    topo = grid.find("Topology")
    assert topo.get("TopologyType") == "Polyvertex" and int(topo.get("NumberOfElements")) == n
    items = [(None, grid.find("Geometry/DataItem"))] + [(a, a.find("DataItem")) for a in grid.findall("Attribute")]
    assert len(items) == 1 + 26
    for attr, di in items:
        path, dset = di.text.strip().rsplit(":/", 1)
        assert path == key + ".hdf5"
        if attr is not None:
            assert attr.get("Name") == dset
        dims = [int(x) for x in di.get("Dimensions").split()]
        assert int(np.prod(dims)) == ds[dset][0][0], dset
        assert di.get("NumberType") == ("Int" if ds[dset][1].dtype == np.int32 else "Float")


def test_csv_matches_ostream(tmp_path):
    n = 1000
    rng = np.random.default_rng(9)
    rows = rng.standard_normal((n, capi.CSV_NCOL)) * 10.0 ** rng.integers(-12, 12, (n, capi.CSV_NCOL))
    rows[:, 3] = rng.integers(0, 5, n)
    rows[0, :] = 0.0
    rows[1, 5] = -0.0
    rows[2, 6:9] = (1e-300, 123456789.0, 0.1)
    exe = str(tmp_path / "csv_ostream")
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(HERE, "cpp", "csv_ostream.cpp"), "-o", exe], check=True)
    rows.tofile(str(tmp_path / "rows.bin"))
    subprocess.run([exe, str(tmp_path / "rows.bin"), str(n), str(tmp_path / "ref.csv")], check=True)
    ref = open(tmp_path / "ref.csv", "rb").read()
    for threads in (1, 3, 8):
        output.write_csv_rows(str(tmp_path / f"t{threads}"), rows, threads=threads)
        assert open(tmp_path / f"t{threads}.csv", "rb").read() == ref, threads


def test_writer_errors(tmp_path):
    with pytest.raises(output.OutputError):
        output.write_csv_rows(str(tmp_path / "nodir" / "x"), np.zeros((2, capi.CSV_NCOL)))
    if output.has_hdf5():
        with pytest.raises(output.OutputError):
            output.write_xdmf_table(str(tmp_path / "nodir" / "x"), 2, np.zeros(64 * 2, np.float32),
                                    np.zeros(6, np.int32))


def test_io_library_exports():
    import ctypes
    import re
    L = ctypes.CDLL(output.IO_LIB_PATH)
    decl = re.findall(r"^\w[\w\s\*]*?\b(wfio_\w+)\(", open(os.path.join(build.ROOT, "include", "weldform_io.h")).read(),
                      re.M)
    assert decl
    for name in decl:
        assert hasattr(L, name), name
