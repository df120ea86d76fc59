"""Host-side checks of libweldform_hip.so that need no GPU: the library loads,
exports every entry point include/weldform_hip.h declares, its structs have
the layout the ctypes binding assumes, and its host-only helpers are right.
No device computation is launched here."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po
from weldform_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "weldform_hip.h")


def declared_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char \*)\s*(wf_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(capi.EXPORTS)
    assert L.wf_abi_version() == 1


def test_struct_layouts_match_ctypes():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "weldform_hip.h"
int main(void) {
  printf("%zu %zu %zu\n", sizeof(wf_config), offsetof(wf_config, cellfac), offsetof(wf_config, dom_min));
  printf("%zu %zu %zu\n", sizeof(wf_material), offsetof(wf_material, E), offsetof(wf_material, th_exp));
  printf("%zu %zu %zu\n", sizeof(wf_scalars), offsetof(wf_scalars, kin_energy_sum), offsetof(wf_scalars, min_T));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as t:
        c = os.path.join(t, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(t, "l")
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    got = list(map(int, out))
    want = [ctypes.sizeof(capi.Config), capi.Config.cellfac.offset, capi.Config.dom_min.offset,
            ctypes.sizeof(capi.Material), capi.Material.E.offset, capi.Material.th_exp.offset,
            ctypes.sizeof(capi.Scalars), capi.Scalars.kin_energy_sum.offset, capi.Scalars.min_T.offset]
    assert got == want


@pytest.mark.parametrize("R,L,r", [(7.5e-3, 30e-3, 0.5e-3), (30e-3, 60e-3, 0.5537e-3 / 2), (4e-3, 9e-3, 0.21e-3)])
def test_generator_matches_oracle_bitwise(R, L, r):
    a = capi.gen_cylinder([0.1, -0.2, 0.05], R, L, r)
    b = po.gen_cylinder([0.1, -0.2, 0.05], R, L, r)
    assert a.shape == b.shape and np.array_equal(a, b)


def test_generator_count_config2():
    assert len(capi.gen_cylinder([0, 0, 0], 30e-3, 60e-3, 0.5537e-3 / 2)) == 976320


def test_errors_without_device_are_codes_not_crashes():
    if capi.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(capi.WFError) as e:
        capi.Solver(10)
    assert e.value.code == capi.EHIP
    assert capi.lib().wf_last_error(None)
    # argument validation happens before any device call
    cfg = capi.Config(precision=7, max_nb=64, cellfac=2.0, ts_nb_inc=5)
    h = ctypes.c_void_p()
    assert capi.lib().wf_create(ctypes.byref(cfg), 10, ctypes.byref(h)) == capi.EINVAL
    cfg = capi.Config(precision=0, max_nb=500, cellfac=2.0, ts_nb_inc=5)
    assert capi.lib().wf_create(ctypes.byref(cfg), 10, ctypes.byref(h)) == capi.EINVAL
    assert capi.lib().wf_step(None, 0, 1) == capi.EINVAL


def test_cases_reference_parameters():
    from weldform_amd import cases
    c = cases.config1()
    assert c.n == 4524
    # Source/WeldForm.cpp:338-358 parameter derivation
    m = c.materials[0]
    assert m["Cs"] == pytest.approx(4611.4, rel=1e-4)
    assert c.dt == pytest.approx(0.4 * 1.2e-3 / m["Cs"], rel=1e-15)
    assert m["Ep"] == pytest.approx(68.9e9 * 1.5e9 / (68.9e9 - 1.5e9), rel=1e-15)
    assert (c.ids == 2).sum() == 156 and (c.ids == 3).sum() == 156
    assert c.dom_max[0] == 30e-3 and c.dom_min[0] == -30e-3
