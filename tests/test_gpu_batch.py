"""Batched ensembles (several cylinder specimens advanced as one particle
set, each with its own material, friction coefficient and pair of rigid
patches, no pair across specimens) against the CPU oracle and against the
same specimens run one by one.

The reference's neighbour-rebuild schedule is global: once any particle's
plastic strain passes 1e-6, every particle's list is rebuilt every
ts_nb_inc steps (Source/SolverLeapfrog.cpp:160-201).  In a batch the first
specimen to yield sets the schedule for all; the oracle, run on the whole
batch, follows the batch's schedule.

The Wang penalty contact (kij = 2 m / dt^2 with static friction,
Source/Contact_Wang.cpp:15-348) makes the run chaotic once the patches
touch (step 12 here): a 1e-15 relative change of the oracle's own initial
positions grows to ~3e-10 of the displacement at step 20 and to a few
per cent by step 50.  So the bars are:
  * step 20 (every top patch in contact): FP64 displacement, plastic strain
    and the per-specimen contact forces of both patches (every step so far)
    within 1e-8 normwise, per specimen; mixed within max(1e-4, 10x the
    oracle's sensitivity to a 1e-7 perturbation);
  * step 120: displacement and per-specimen force histories within 10x
    the oracle's sensitivity (1e-15 for FP64, 1e-7 for mixed), and each
    specimen's mean top-patch force over steps 61-120 within 3 % or 3x its
    sensitivity;
  * per-specimen sums add up to the patch totals (1e-12 of the largest);
  * a specimen alone against its share of the batch at step 20: within
    1e-8 (FP64) or 1e-4 (mixed), or 10x the oracle's sensitivity, since a
    specimen alone follows its own neighbour-rebuild schedule."""
import numpy as np
import pytest

from oracle_case import normwise_err, oracle_from_case
from weldform_amd import autotune, capi, cases

pytestmark = pytest.mark.gpu

STEPS = 120
EARLY = 20
NSPEC = 6
MEAN_FROM = 60


def _batch_case():
    return cases.config4(n_spec=NSPEC, n_range=(1500, 5000), steps=STEPS, v_top=-20.0, seed=5)


def _oracle(c, eps=0.0):
    d = oracle_from_case(c, threads=8)
    if eps:
        rng = np.random.default_rng(5)
        d.f("X")[:c.n] = c.x * (1 + eps * rng.uniform(-1, 1, size=c.x.shape))
    d.solve_init(c.dt)
    G, early = [], None
    for s in range(STEPS):
        d.step()
        G.append([[d.group_contact_force(g, m) for m in range(2)] for g in range(NSPEC)])
        if s + 1 == EARLY:
            early = (d.f("DISP")[:c.n].copy(), d.f("PLSTRAIN")[:c.n].copy())
    return d, np.array(G), early  # G: (steps, group, mesh, 3)


def _errs(c, disp, G, early_disp, o_disp, o_G, o_early):
    """Per-specimen errors: step-20 displacement and force histories, final
    displacement, force histories (z) and mean top force."""
    out = []
    for i in range(NSPEC):
        sel = c.material == i
        out.append({
            "disp_early": normwise_err(early_disp[sel], o_early[sel]),
            "F_early": normwise_err(G[:EARLY, i], o_G[:EARLY, i]),
            "disp": normwise_err(disp[sel], o_disp[sel]),
            "F_z": normwise_err(G[:, i, :, 2], o_G[:, i, :, 2]),
            "mean_top": abs(G[MEAN_FROM:, i, 0, 2].mean() / o_G[MEAN_FROM:, i, 0, 2].mean() - 1)})
    return out


@pytest.fixture(scope="module")
def batch():
    c = _batch_case()
    d, G, early = _oracle(c)
    o_disp = d.f("DISP")[:c.n].copy()
    sens = {}
    for eps in (1e-15, 1e-7):
        dp, Gp, ep = _oracle(c, eps)
        sens[eps] = _errs(c, dp.f("DISP")[:c.n], Gp, ep[0], o_disp, G, early[0])
    return c, d, G, early, sens


def _gpu(c, precision):
    g = cases.make_solver(c, precision=precision)
    g.step(EARLY)
    g.sync()
    early = (g.get("disp"), g.get("pl_strain"))
    g.step(STEPS - EARLY)
    g.sync()
    G = np.stack([np.stack([g.group_history(1, i, m) for m in range(2)], axis=1) for i in range(NSPEC)], axis=1)
    tot = np.stack([g.contact_force_history(m) for m in range(2)], axis=1)
    disp = g.get("disp")
    g.close()
    return disp, G, early, tot


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_batch_matches_oracle(batch, precision):
    c, d, G, early, sens_all = batch
    fp64 = precision == capi.FP64
    print(f"batch: n={c.n} specimens={NSPEC} steps={STEPS} precision={precision}", flush=True)
    assert np.abs(d.f("PLSTRAIN")).max() > 0
    assert (np.abs(G[:EARLY, :, 0, 2]).max(axis=0) > 0).all()  # every top patch touches early
    assert (np.abs(G[:, :, 1, 2]).max(axis=0) > 0).all()  # and every bottom patch later
    disp, Gg, (e_disp, e_pl), tot = _gpu(c, precision)
    assert Gg.shape == G.shape
    # the per-specimen sums add up to each patch mesh's total
    assert np.abs(Gg.sum(axis=1) - tot).max() < 1e-12 * np.abs(tot).max()
    errs = _errs(c, disp, Gg, e_disp, d.f("DISP")[:c.n], G, early[0])
    sens = sens_all[1e-15 if fp64 else 1e-7]
    floor = 1e-8 if fp64 else 1e-4
    for i in range(NSPEC):
        sel = c.material == i
        e_pl_i = normwise_err(e_pl[sel], early[1][sel])
        print(f"specimen {i}: n={int(sel.sum())} mu={c.meta['friction'][i]:.3f} errors "
              + " ".join(f"{k} {v:.2e}" for k, v in errs[i].items()) + f" pl_early {e_pl_i:.2e}; sensitivity "
              + " ".join(f"{k} {v:.2e}" for k, v in sens[i].items()), flush=True)
        if fp64:
            assert errs[i]["disp_early"] < 1e-8 and errs[i]["F_early"] < 1e-8 and e_pl_i < 1e-8
        else:
            for k in ("disp_early", "F_early"):
                assert errs[i][k] < max(floor, 10 * sens[i][k]), k
        for k in ("disp", "F_z"):
            assert errs[i][k] < max(floor, 10 * sens[i][k]), k
        assert errs[i]["mean_top"] < max(0.03, 3 * sens[i]["mean_top"])


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_lone_specimens_match_their_share(batch, precision):
    c, _, _, _, sens_all = batch
    fp64 = precision == capi.FP64
    sens = sens_all[1e-15 if fp64 else 1e-7]
    floor = 1e-8 if fp64 else 1e-4
    g = cases.make_solver(c, precision=precision)
    g.step(EARLY)
    g.sync()
    disp = g.get("disp")
    G = np.stack([np.stack([g.group_history(1, i, m) for m in range(2)], axis=1) for i in range(NSPEC)], axis=1)
    g.close()
    for i in range(NSPEC):
        s = cases.specimen(c, i)
        lone = cases.make_solver(s, precision=precision)
        lone.step(EARLY)
        lone.sync()
        F = np.stack([lone.contact_force_history(m) for m in range(2)], axis=1)
        e_d = normwise_err(disp[c.material == i], lone.get("disp"))
        e_f = normwise_err(G[:, i], F)
        lone.close()
        print(f"specimen {i} alone (precision {precision}): disp {e_d:.2e} forces {e_f:.2e}", flush=True)
        assert e_d < max(floor, 10 * sens[i]["disp_early"]) and e_f < max(floor, 10 * sens[i]["F_early"])


def test_autotune_shapes_agree(batch):
    c = batch[0]
    best, table = autotune.autotune(c, precision=capi.MIXED, steps=5, warmup=2, lanes=(2,), stages=(-1, 512, 0))
    print("autotune", best, table, flush=True)
    assert best in [{"lanes": r["lanes"], "stage_max": r["stage_max"]} for r in table]
    out = []
    for st in (-1, 512, 0):
        g = cases.make_solver(c, precision=capi.MIXED, lanes=2, stage_max=st)
        g.step(10)
        g.sync()
        out.append(g.get("disp"))
        g.close()
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
