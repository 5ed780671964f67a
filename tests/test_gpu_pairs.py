"""GPU: the FP32 inner CG as pairs of iterations (k_cgp: two sweeps per two
iterations, s-step CG with s = 2) against the fused one-kernel iteration and
the FP64 oracle.

Both run CG on the correction up to FP32 rounding: the pair form takes its
scalars from six sums of the basis p, A p, A^2 p, r, A r instead of from the
vectors each iteration forms.  Each form must meet the FP64 true residual of
1e-10 (checked by kernels independent of the solver), agree with the
oracle's FP64 solution within 1e-9 in every field-summary quantity, and need
about the same number of iterations (an operator wrong at a wave, block or
grid edge, or a wrong pair scalar, would converge far more slowly or not at
all, since the FP64 refresh measures the true residual).  The grids are the
fused kernel's: one lane, one wave, ragged last blocks, a single row band
and bands shorter than the 2-row halo."""
import numpy as np
import pytest

import oracle as O
import tealeaf_hip as T
from tealeaf_hip import _lib, decks

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b)))


def _deck(nx, ny, seed, steps, extra=""):
    return decks.random_rect_deck(seed, n=nx, steps=steps, nrect=8, precision="fp32_ir").replace(
        f"y_cells={nx}", f"y_cells={ny}") + extra


def _oracle(text, steps):
    d = O.parse_deck(text)
    d.stop_rule, d.rel_tol, d.maxiters = O.STOP_RELATIVE, 1e-12, 200000
    o = O.Oracle(d)
    o.initialise()
    return o.run(steps)[0]


def _run(text, steps, form, **kw):
    s = T.Settings.from_text(text).set(fp32_kernels=form, **kw)
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, steps)
    u = c.field("u")
    c.close()
    return summ, st, u


@pytest.mark.parametrize("nx,ny", [(256, 256), (64, 40), (8, 12), (4, 4), (1028, 9), (1032, 70), (2048, 5)])
def test_pairs_match_oracle_and_fused_form(nx, ny):
    steps = 2
    print(f"case: FP32 CG pairs vs fused, {nx}x{ny}, {steps} steps", flush=True)
    text = _deck(nx, ny, 41, steps)
    osum = _oracle(text, steps)
    ps, pst, _ = _run(text, steps, _lib.TL_FP32_KERNELS_PAIRS)
    fs, fst, _ = _run(text, steps, _lib.TL_FP32_KERNELS_FUSED)
    for t in range(steps):
        for name, summ, st in (("pairs", ps, pst), ("fused", fs, fst)):
            assert st[t]["converged"] == 1 and st[t]["rel_resid"] <= 1e-10, (name, t, st[t])
            assert _rel(summ[t], osum[t]) < 1e-9, (name, t, summ[t], osum[t])
        pi, fi = pst[t]["iterations"], fst[t]["iterations"]
        print(f"  step {t + 1}: pairs {pi} ({pst[t]['refreshes']} refreshes), fused {fi} "
              f"({fst[t]['refreshes']})", flush=True)
        assert abs(pi - fi) <= max(6, 0.05 * fi), (t, pi, fi)
        # the pair tests the refresh rule after its first step as well, so it
        # refreshes as often as the one-kernel form (the second step of a pair
        # that reaches a refresh is a null step)
        assert pst[t]["refreshes"] == fst[t]["refreshes"], (t, pst[t]["refreshes"], fst[t]["refreshes"])


def test_pairs_validation_and_iteration_limit():
    print("case: pairs form on nx % 4 != 0 (refused) and an odd iteration limit", flush=True)
    text = _deck(66, 20, 3, 1)
    s = T.Settings.from_text(text).set(fp32_kernels=_lib.TL_FP32_KERNELS_PAIRS)
    c = T.initialiseapp(s)
    with pytest.raises(T.TeaLeafError):
        T.step(c, s)
    c.close()
    # a limit of 9 iterations stops both forms after exactly 9, unconverged,
    # with the partial correction kept: the true residual is below the
    # starting one (the same for an even limit, 10).  The two forms refresh at
    # different iterations and CG's residual norm is not monotone, so their
    # residuals after so few steps are not compared with each other.
    text = _deck(128, 96, 5, 1)
    for lim in (9, 10):
        for name, form in (("pairs", _lib.TL_FP32_KERNELS_PAIRS), ("fused", _lib.TL_FP32_KERNELS_FUSED)):
            summ, st, u = _run(text, 1, form, max_iters=lim)
            print(f"  max_iters {lim}: {name} {st[0]['iterations']} iterations, converged {st[0]['converged']}, "
                  f"rel_resid {st[0]['rel_resid']:.3e} (start {st[0]['init_rel_resid']:.3e})", flush=True)
            assert st[0]["iterations"] == lim and st[0]["converged"] == 0, (name, lim, st[0])
            assert st[0]["rel_resid"] < st[0]["init_rel_resid"], (name, lim, st[0])


def test_pairs_deterministic():
    print("case: pairs form run twice bit for bit, 256x192, 2 steps", flush=True)
    text = _deck(256, 192, 7, 2)
    a, ast_, ua = _run(text, 2, _lib.TL_FP32_KERNELS_PAIRS)
    b, bst, ub = _run(text, 2, _lib.TL_FP32_KERNELS_PAIRS)
    assert np.array_equal(a, b) and np.array_equal(ua, ub)
    assert [x["iterations"] for x in ast_] == [x["iterations"] for x in bst]
    assert all(x["iterations"] % 2 == 0 for x in ast_)
