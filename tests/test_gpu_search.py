"""The neighbour search on grids the headline case does not produce, against
a brute-force neighbour set (bit-exact, as for the other search tests):

- overfull cells: six copies of the config-1 billet, slightly offset, each
  of its own material.  Pairs form only within a material (Source/
  Neighbour.cpp:400-414), so every row keeps ~55 neighbours while the
  search cells hold ~80 particles: the cells of more than 64 particles are
  sorted by the first lane of their wavefront;
- a sparse grid: two billets half a metre apart, so the cell-start scan runs
  over many chunks of mostly empty cells and the per-cell pre-pass handles
  them."""
import numpy as np
import pytest

from oracle_case import brute_pairs
from weldform_amd import capi, cases

pytestmark = pytest.mark.gpu


def _merge(parts):
    c = parts[0]
    out = cases.Case(name="merged", x=np.concatenate([p.x for p in parts]), h=np.concatenate([p.h for p in parts]),
                     mass=np.concatenate([p.mass for p in parts]),
                     density=np.concatenate([p.density for p in parts]),
                     ids=np.concatenate([p.ids for p in parts]),
                     material=np.concatenate([np.full(p.n, k, dtype=np.int32) for k, p in enumerate(parts)]),
                     materials={k: dict(c.materials[0]) for k in range(len(parts))},
                     sigmay=np.concatenate([p.sigmay for p in parts]), dt=c.dt, steps=1,
                     bcs=c.bcs, dom_max=c.dom_max, dom_min=c.dom_min)
    return out


def _brute_by_material(case):
    pairs = []
    for k in np.unique(case.material):
        idx = np.nonzero(case.material == k)[0]
        p = brute_pairs(case.x[idx], case.h[idx])
        pairs.append(idx[p])
    allp = np.concatenate(pairs)
    allp = np.sort(allp, axis=1)
    return allp[np.lexsort((allp[:, 1], allp[:, 0]))]


def _sorted_pairs(g):
    p = np.sort(g.astype(np.int64), axis=1)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def _max_per_cell(case, grid):
    hmax = case.h.max()
    lo = case.x.min(0) - hmax / 2
    hi = case.x.max(0) + hmax / 2
    size = (hi - lo) / grid
    c = np.clip(np.floor((case.x - lo) / size).astype(np.int64), 0, grid - 1)
    key = (c[:, 0] * grid[1] + c[:, 1]) * grid[2] + c[:, 2]
    return np.bincount(key).max(), int(np.prod(grid))


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_overfull_cells(precision):
    parts = []
    for k in range(6):
        p = cases.config1(steps=1)
        p.x = p.x + np.array([0.11, 0.07, 0.05]) * k * 1e-3
        p.dom_max, p.dom_min = (-1e11, -1e11, -1e11), (1e11, 1e11, 1e11)
        parts.append(p)
    c = _merge(parts)
    s = cases.make_solver(c, precision=precision)
    s.rebuild_neighbours()
    ref, grid = s.cell_grids()
    most, ncells = _max_per_cell(c, np.asarray(grid))
    print(f"overfull cells: n={c.n} search grid {grid} most particles in a cell {most}", flush=True)
    assert most > 64
    assert np.array_equal(_sorted_pairs(s.neighbour_pairs()), _brute_by_material(c))
    s.close()


def test_sparse_grid():
    a = cases.config1(steps=1)
    b = cases.config1(steps=1)
    b.x = b.x + np.array([0.5, 0.0, 0.0])
    for p in (a, b):
        p.dom_max, p.dom_min = (-1e11, -1e11, -1e11), (1e11, 1e11, 1e11)
    c = _merge([a, b])
    c.material[:] = 0
    c.materials = {0: dict(a.materials[0])}
    s = cases.make_solver(c)
    s.rebuild_neighbours()
    ref, grid = s.cell_grids()
    ncells = int(np.prod(grid))
    print(f"sparse grid: n={c.n} search grid {grid} ({ncells} cells)", flush=True)
    assert ncells > 8 * 1024
    assert np.array_equal(_sorted_pairs(s.neighbour_pairs()), _brute_by_material(c))
    s.close()
