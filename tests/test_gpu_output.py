"""Output tables packed on the device (wf_pack_xdmf / wf_pack_csv, the
particle tables of Domain::WriteXDMF and Domain::WriteCSV,
Source/Output.cpp:75-554) against the same tables packed on the host from
wf_get_field / wf_get_int_field of the same state.  The packing is a
permutation to the caller's order, a cast to float32 and Sigma_eq formed
with the same FP64 operations, so the bar is bit-exact.  The run has
contact under way (free-surface IDs, contact-pair counts and contact
forces live), in FP64 and mixed precision; then the files written from the
device tables read back equal to the tables."""
import numpy as np
import pytest

import h5read
from weldform_amd import capi, cases, output

pytestmark = pytest.mark.gpu


def _host_fields(g):
    fl = {k: g.get(k) for k in ("x", "v", "a", "disp", "pressure", "density", "mass", "h", "sigma", "shear", "strain",
                                "strain_rate", "strain_pl", "T", "pl_strain", "ps_energy", "contforce")}
    fl.update({"id": g.get_int("id"), "nb": g.get_int("nb"), "contnb": g.get_int("contnb")})
    return fl


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_device_tables_match_host_packing(precision, tmp_path):
    c = cases.contact_small(steps=300)
    print(f"output tables: n={c.n} precision={precision}", flush=True)
    g = cases.make_solver(c, precision=precision)
    g.step(300)
    g.sync()
    fl = _host_fields(g)
    assert fl["contnb"].max() > 0 and np.abs(fl["contforce"]).max() > 0 and fl["pl_strain"].max() > 0
    assert (fl["id"] == 1).any()  # free-surface particles
    f_h, iv_h, rows_h = output.table_from_fields(fl)
    f, iv = g.pack_xdmf()
    rows = g.pack_csv()
    assert np.array_equal(iv, iv_h)
    bad = {name: int((a != b).sum()) for (name, a), b in
           zip(output.datasets(f, iv, c.n).items(), output.datasets(f_h, iv_h, c.n).values()) if not np.array_equal(a, b)}
    assert not bad, bad
    assert np.array_equal(rows, rows_h)
    # a second pack after more steps sees the new state
    g.step(5)
    g.sync()
    f2, _ = g.pack_xdmf()
    assert np.array_equal(output.datasets(f2, iv, c.n)["Position"], g.get("x").astype(np.float32).reshape(-1))
    # files written from the device tables
    key = str(tmp_path / "contact_0001")
    output.write_csv(g, key)
    lines = open(key + ".csv").read().splitlines()
    assert lines[0] == output.CSV_HEADER and len(lines) == c.n + 1
    if output.has_hdf5():
        output.write_xdmf(g, key)
        ds = h5read.read(key + ".hdf5")
        want = output.datasets(*g.pack_xdmf(), c.n)
        for name, arr in want.items():
            assert np.array_equal(ds[name][1], arr), name
        assert ds["NP"][1][0] == c.n
    g.close()


def test_pack_before_init_fails():
    c = cases.small_jc(steps=1)
    s = capi.Solver(c.n)
    with pytest.raises(capi.WFError):
        s.pack_xdmf()
    with pytest.raises(capi.WFError):
        s.pack_csv()
    s.close()
