"""GPU: two solver steps per sweep against one per sweep: PPCG's inner
Chebyshev steps (k_ppcg2 against k_vs64, PPCG.jl:78-83) and the Chebyshev
solver's iterations (k_cheby2 against k_vs64, Cheby.jl:93-107).

The two-step kernels apply the single step's operations in the same order,
so every field and statistic must be bit-identical after whole timesteps:
u, r, sd, p, w over the whole allocation (halos included), the iteration
counts, the error and the eigenvalue bounds.  Grids cover several waves per
row (the columns next to a wave that the edge lanes compute themselves), a
ragged last wave, an odd column count, bands shorter than the kernel's
2-row reach and a single row of waves; inner step counts cover pairs only,
pairs plus an odd last single step, and too few steps for a pair."""
import numpy as np
import pytest

import tealeaf_hip as T
from tealeaf_hip import _lib, decks

pytestmark = pytest.mark.gpu

FIELDS = ("u", "r", "sd", "p", "w", "u0")


def _deck(nx, ny, steps, inner, solver="ppcg"):
    return decks.reference_deck(nx, steps, solver=solver, presteps=8, ppcg_inner_steps=inner,
                                max_iters=4000).replace(f"y_cells={nx}", f"y_cells={ny}")


def _run(text, steps, pairs, key="ppcg_pairs"):
    s = T.Settings.from_text(text).set(**{key: pairs})
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, steps)
    fields = {f: c.field(f) for f in FIELDS}
    c.close()
    return summ, st, fields


@pytest.mark.parametrize("nx,ny,inner", [(128, 128, 10), (300, 70, 10), (257, 33, 5), (130, 9, 4),
                                         (520, 40, 7), (64, 5, 6), (96, 64, 3)])
def test_pairs_bit_identical_to_single_steps(nx, ny, inner):
    steps = 2
    print(f"case: PPCG two inner steps per sweep vs one, {nx}x{ny}, {inner} inner steps, {steps} steps",
          flush=True)
    text = _deck(nx, ny, steps, inner)
    a_sum, a_st, a_f = _run(text, steps, _lib.TL_PPCG_PAIRS_AUTO)
    b_sum, b_st, b_f = _run(text, steps, _lib.TL_PPCG_PAIRS_OFF)
    assert np.array_equal(a_sum, b_sum), (a_sum, b_sum)
    for t in range(steps):
        for k in ("iterations", "cg_iters", "cheby_iters", "error", "eigmin", "eigmax", "converged"):
            assert a_st[t][k] == b_st[t][k], (t, k, a_st[t][k], b_st[t][k])
    for f in FIELDS:
        assert np.array_equal(a_f[f], b_f[f]), f
    print(f"  iterations {[x['iterations'] for x in a_st]}, error {[x['error'] for x in a_st]}", flush=True)


@pytest.mark.parametrize("nx,ny", [(128, 128), (300, 70), (257, 33), (130, 9), (64, 5), (512, 96)])
def test_cheby_pairs_bit_identical_to_single_iterations(nx, ny):
    steps = 2
    print(f"case: Chebyshev two iterations per sweep vs one, {nx}x{ny}, {steps} steps", flush=True)
    text = _deck(nx, ny, steps, 10, solver="cheby")
    a_sum, a_st, a_f = _run(text, steps, _lib.TL_CHEBY_PAIRS_AUTO, "cheby_pairs")
    b_sum, b_st, b_f = _run(text, steps, _lib.TL_CHEBY_PAIRS_OFF, "cheby_pairs")
    assert np.array_equal(a_sum, b_sum), (a_sum, b_sum)
    for t in range(steps):
        for k in ("iterations", "cg_iters", "cheby_iters", "est_iters", "error", "eigmin", "eigmax", "converged"):
            assert a_st[t][k] == b_st[t][k], (t, k, a_st[t][k], b_st[t][k])
        if nx >= 128 and ny >= 64:
            assert a_st[t]["cheby_iters"] > 20, a_st[t]  # the Chebyshev phase ran, not only CG
    for f in FIELDS:
        d = a_f[f] != b_f[f]
        assert not d.any(), (f, "cells differing (interior, halo):", int(d[2:-2, 2:-2].sum()),
                             int(d.sum() - d[2:-2, 2:-2].sum()))
    print(f"  iterations {[x['iterations'] for x in a_st]} (Chebyshev {[x['cheby_iters'] for x in a_st]})",
          flush=True)


def test_pairs_validation():
    print("case: ppcg_pairs / cheby_pairs validation", flush=True)
    for key, solver in (("ppcg_pairs", "ppcg"), ("cheby_pairs", "cheby")):
        s = T.Settings.from_text(_deck(64, 64, 1, 10, solver)).set(**{key: 7})
        c = T.initialiseapp(s)
        with pytest.raises(T.TeaLeafError):
            T.step(c, s)
        c.close()
