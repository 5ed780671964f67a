"""GPU: the learned-warm-start configurations and the HBM-capacity probe.

* C4: the committed warm-start network (trained on 1024^2 trajectories,
  seeds 0-199) on an unseen seed at 4096^2, FP64 CG: every solve meets the
  1e-10 relative residual, the summaries agree with the run without the
  network, and the acceptance count and iteration difference are reported.
* C5: the learned start with the FP32-refined CG at 8192^2 (the benchmark's
  method) on a held-out seed: 1e-10 every step, summaries within 1e-9 of
  plain FP64 CG from the previous-step start.
* 16384^2: one FP32-refined timestep (about 40 GB of fields resident):
  residual, vol and mass exact, ie conserved.
"""
import os

import numpy as np
import pytest

import tealeaf_hip as T
from tealeaf_hip import _lib, decks, unet

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "tealeaf.jl_amd", "weights", "unet_warmstart.npy")


def _rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b)) / np.abs(np.asarray(b)))


def _run(text, steps, weights=None, every=False):
    s = T.Settings.from_text(text)
    if every:  # the network evaluated at every solve, not only after acceptances
        s.set(warm_policy=_lib.TL_WARM_POLICY_EVERY)
    c = T.initialiseapp(s)
    if weights is not None:
        unet.load_weights(c, weights)
    out = []
    for _ in range(steps):
        st = T.step(c, s)
        out.append((T.fieldsummary(c), st))
    c.close()
    return out


@pytest.mark.timeout(600)
def test_c4_learned_start_4096_unseen_seed():
    steps = 3
    print(f"case: C4 learned start, FP64 CG, 4096^2 seed 1000, {steps} steps", flush=True)
    base_text = decks.random_rect_deck(1000, n=4096, steps=steps)
    base = _run(base_text, steps)
    learn = _run(base_text.replace("warm_start=none", "warm_start=learned"), steps, WEIGHTS, every=True)
    accepted = 0
    for t, ((sb, stb), (sl, stl)) in enumerate(zip(base, learn)):
        assert stb["converged"] == 1 and stb["rel_resid"] <= 1e-10, stb
        assert stl["converged"] == 1 and stl["rel_resid"] <= 1e-10, stl
        assert _rel(sl, sb) < 1e-9, (t, sl, sb)
        accepted += stl["start"] == _lib.TL_START_WARM
        assert stl["fallback"] == (stl["start"] != _lib.TL_START_WARM)
        print(f"  step {t + 1}: iterations {stb['iterations']} without / {stl['iterations']} with the network "
              f"({'accepted' if stl['start'] == _lib.TL_START_WARM else 'rejected'}), solve "
              f"{stb['solve_ms']:.0f} / {stl['solve_ms']:.0f} ms, initial residual {stb['init_rel_resid']:.3e} / "
              f"{stl['init_rel_resid']:.3e}", flush=True)
    print(f"  accepted {accepted} of {steps}", flush=True)


@pytest.mark.timeout(600)
def test_c5_learned_fp32_refined_8192():
    steps = 2
    print(f"case: C5 learned start + FP32-refined CG vs FP64 CG, 8192^2 seed 1002, {steps} steps", flush=True)
    flag = _run(decks.random_rect_deck(1002, n=8192, steps=steps, precision="fp32_ir", warm_start="learned"),
                steps, WEIGHTS)
    base = _run(decks.random_rect_deck(1002, n=8192, steps=steps), steps)
    for t, ((sf, stf), (sb, stb)) in enumerate(zip(flag, base)):
        assert stf["converged"] == 1 and stf["rel_resid"] <= 1e-10, stf
        assert stb["converged"] == 1 and stb["rel_resid"] <= 1e-10, stb
        assert _rel(sf, sb) < 1e-9, (t, sf, sb)
        print(f"  step {t + 1}: FP32-refined {stf['iterations']} iterations ({stf['refreshes']} refreshes, "
              f"start {'learned' if stf['start'] == _lib.TL_START_WARM else 'previous'}) {stf['solve_ms']:.0f} ms; "
              f"FP64 {stb['iterations']} iterations {stb['solve_ms']:.0f} ms; max rel diff {_rel(sf, sb):.2e}",
              flush=True)


@pytest.mark.timeout(900)
def test_16384_fp32_refined_one_step_properties():
    print("case: 16384^2 random rectangles seed 1003, 1 FP32-refined step", flush=True)
    s = T.Settings.from_text(decks.random_rect_deck(1003, n=16384, steps=1, precision="fp32_ir"))
    c = T.initialiseapp(s)
    before = T.fieldsummary(c)
    st = T.step(c, s)
    after = T.fieldsummary(c)
    c.close()
    print(f"  {st['iterations']} iterations, {st['refreshes']} refreshes, {st['solve_ms'] / 1e3:.1f} s, "
          f"rel. residual {st['rel_resid']:.3e}", flush=True)
    assert st["converged"] == 1 and st["rel_resid"] <= 1e-10, st
    assert after[0] == before[0] and abs(after[1] / before[1] - 1) < 1e-13
    assert abs(after[2] / before[2] - 1) < 1e-8, (before, after)
