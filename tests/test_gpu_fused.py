"""GPU: the FP32 inner CG as one kernel per iteration (k_cgf, single
reduction) against the two-kernel form (k_cgg) and the FP64 oracle.

Both forms run the same algorithm up to FP32 rounding: each must meet the
FP64 true residual of 1e-10 (checked by kernels independent of the solver),
agree with the oracle's FP64 solution within 1e-9 in every field-summary
quantity, and need about the same number of iterations (a kernel that
applied a wrong operator at a wave, block or grid edge would converge much
more slowly, or not at all, since the FP64 refresh measures the true
residual).  Grids cover one lane, one wave, ragged last blocks, a single
row band and bands shorter than the kernel's 2-row halo."""
import numpy as np
import pytest

import oracle as O
import tealeaf_hip as T
from tealeaf_hip import _lib, decks

pytestmark = pytest.mark.gpu

REL = "stop_rule=relative\nrel_tol=1e-10\nprecision=fp32_ir\n"


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b)))


def _deck(nx, ny, seed, steps):
    return decks.random_rect_deck(seed, n=nx, steps=steps, nrect=8, precision="fp32_ir").replace(
        f"y_cells={nx}", f"y_cells={ny}")


def _oracle(text, steps):
    d = O.parse_deck(text)
    d.stop_rule, d.rel_tol, d.maxiters = O.STOP_RELATIVE, 1e-12, 200000
    o = O.Oracle(d)
    o.initialise()
    return o.run(steps)[0]


def _run(text, steps, form, launch=T.TL_LAUNCH_STREAM):
    s = T.Settings.from_text(text).set(fp32_kernels=form, launch=launch)
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, steps)
    u = c.field("u")
    c.close()
    return summ, st, u


@pytest.mark.parametrize("nx,ny", [(256, 256), (64, 40), (8, 12), (4, 4), (1028, 9), (1032, 70), (2048, 5)])
def test_fused_matches_oracle_and_two_kernel_form(nx, ny):
    steps = 2
    print(f"case: fused vs two-kernel FP32 CG, {nx}x{ny}, {steps} steps", flush=True)
    text = _deck(nx, ny, 41, steps)
    osum = _oracle(text, steps)
    fs, fst, _ = _run(text, steps, _lib.TL_FP32_KERNELS_FUSED)
    ts, tst, _ = _run(text, steps, _lib.TL_FP32_KERNELS_TWO)
    for t in range(steps):
        for name, summ, st in (("fused", fs, fst), ("two", ts, tst)):
            assert st[t]["converged"] == 1 and st[t]["rel_resid"] <= 1e-10, (name, t, st[t])
            assert _rel(summ[t], osum[t]) < 1e-9, (name, t, summ[t], osum[t])
        fi, ti = fst[t]["iterations"], tst[t]["iterations"]
        print(f"  step {t + 1}: fused {fi} ({fst[t]['refreshes']} refreshes), two-kernel {ti}", flush=True)
        assert abs(fi - ti) <= max(4, 0.05 * ti), (t, fi, ti)


def test_fused_auto_selection_and_validation():
    print("case: form selection (auto) and the fused form's nx % 4 requirement", flush=True)
    # nx % 4 != 0: auto takes the two-kernel form, an explicit fused request is refused
    text = _deck(66, 20, 3, 1)
    s = T.Settings.from_text(text).set(fp32_kernels=_lib.TL_FP32_KERNELS_FUSED)
    c = T.initialiseapp(s)
    with pytest.raises(T.TeaLeafError):
        T.step(c, s)
    c.close()
    a, ast_, _ = _run(text, 1, _lib.TL_FP32_KERNELS_AUTO)
    b, bst, _ = _run(text, 1, _lib.TL_FP32_KERNELS_TWO)
    assert np.array_equal(a, b) and ast_[0]["iterations"] == bst[0]["iterations"]
    # nx % 4 == 0: auto is the pairs of iterations with stream launches and
    # the fused form with hipGraph batches, bit for bit
    text = _deck(128, 96, 3, 1)
    for launch, form in ((T.TL_LAUNCH_STREAM, _lib.TL_FP32_KERNELS_PAIRS), (T.TL_LAUNCH_GRAPH, _lib.TL_FP32_KERNELS_FUSED)):
        a, ast_, ua = _run(text, 1, _lib.TL_FP32_KERNELS_AUTO, launch)
        b, bst, ub = _run(text, 1, form, launch)
        assert np.array_equal(a, b) and np.array_equal(ua, ub) and ast_[0]["iterations"] == bst[0]["iterations"]


def test_fused_graph_replays_bit_identical_to_stream():
    print("case: fused FP32 CG, hipGraph batches vs stream launches, 256x192, 3 steps", flush=True)
    text = _deck(256, 192, 7, 3)
    a, ast_, ua = _run(text, 3, _lib.TL_FP32_KERNELS_FUSED, T.TL_LAUNCH_STREAM)
    b, bst, ub = _run(text, 3, _lib.TL_FP32_KERNELS_FUSED, T.TL_LAUNCH_GRAPH)
    assert np.array_equal(a, b) and np.array_equal(ua, ub)
    for x, y in zip(ast_, bst):
        assert x["iterations"] == y["iterations"] > 64 and x["refreshes"] == y["refreshes"]
