"""wf_pack_xdmf_noalloc (the allocation-free variant of wf_pack_xdmf,
staged in the library's static device scratch) against wf_pack_xdmf on the
same state: bit-exact tables, on two calls in a row (the second context
smaller than the first, so the scratch still holds the first call's larger
table), then a context beyond the scratch's particle limit, which must be
refused with WF_EINVAL while wf_pack_xdmf still serves it."""
import numpy as np
import pytest

from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


def _same_tables(g):
    f_ref, iv_ref = g.pack_xdmf()
    f, iv = g.pack_xdmf_noalloc()
    assert np.array_equal(iv, iv_ref)
    assert np.array_equal(f.view(np.uint32), f_ref.view(np.uint32))
    return f


def test_noalloc_pack_matches_allocating_pack_twice():
    big = cases.contact_small(steps=60)
    small = cases.cylinder_case("cyl_upset_small_jc", 4.0e-3, 8.0e-3, 1.0e-3, 2700.0, 68.9e9, 0.3, cases.JC_AL,
                                steps=20, jitter=1e-3, seed=3, v_top=-5.0, T0=300.0)
    assert small.n < big.n
    print(f"noalloc pack: first n={big.n} (mixed, contact), then n={small.n} (fp64)", flush=True)
    g = cases.make_solver(big, precision=capi.MIXED)
    g.step(60)
    g.sync()
    f1 = _same_tables(g)
    assert np.abs(f1).max() > 0
    g.close()
    h = cases.make_solver(small, precision=capi.FP64)
    h.step(20)
    h.sync()
    f2 = _same_tables(h)
    assert f2.size < f1.size
    h.close()


def test_noalloc_pack_refuses_beyond_scratch():
    nmax = capi.lib().wf_pack_xdmf_noalloc_max()
    assert nmax == (64 << 20) // (4 * (capi.XDMF_NF + capi.XDMF_NI))
    c = cases.cylinder_case("cyl_over_scratch", 15e-3, 30e-3, 0.42e-3, 7850.0, 200e9, 0.29, cases.STEEL_JC, steps=1,
                            T0=300.0)
    assert c.n > nmax
    print(f"noalloc pack: n={c.n} > max {nmax}", flush=True)
    g = cases.make_solver(c, precision=capi.MIXED)
    g.sync()
    with pytest.raises(capi.WFError) as e:
        g.pack_xdmf_noalloc()
    assert e.value.code == capi.EINVAL and "use wf_pack_xdmf" in str(e.value)
    f, iv = g.pack_xdmf()
    assert f.size == capi.XDMF_NF * c.n
    g.close()
