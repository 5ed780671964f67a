"""hipGraph batches of CG iterations (TL_LAUNCH_GRAPH) against stream launches.

A replay runs the same kernels in the same order on the same buffers as the
stream launches; only the iteration number comes from device memory.  Every
field and every solver statistic must therefore be bit-identical to the
stream run, after every timestep: each solve replays the cached graph many
times, with the stream's own work (FP64 refreshes, finalise, halo updates,
field summary) and host synchronisations between the replays.  The FP32
cases name their form: the automatic choice differs between the two launch
modes (pairs of iterations run as stream launches only).
"""
import json
import os

import numpy as np
import pytest

import tealeaf_hip as T
from tealeaf_hip import decks

pytestmark = pytest.mark.gpu

FIELDS = ("density", "energy", "u", "u0", "p", "r", "w", "kx", "ky")
STATS = ("iterations", "cg_iters", "inner_iters", "refreshes", "converged", "fallback", "error", "rel_resid",
         "init_rel_resid", "start")


def _run(text, launch, steps, dts=None):
    s = T.Settings.from_text(text).set(launch=launch)
    c = T.initialiseapp(s)
    out = []
    for t in range(steps):
        if dts:
            s._deck.dt_init = dts[t]
        st = T.step(c, s)
        out.append((T.fieldsummary(c), st, {f: c.field(f) for f in FIELDS}))
    return out


def _same(a, b):
    for t, ((sa, ta, fa), (sb, tb, fb)) in enumerate(zip(a, b)):
        assert np.array_equal(sa, sb), (t, sa, sb)
        for k in STATS:
            assert ta[k] == tb[k], (t, k, ta[k], tb[k])
        for f in FIELDS:
            assert np.array_equal(fa[f], fb[f]), (t, f, np.max(np.abs(fa[f] - fb[f])))


REL = "stop_rule=relative\nrel_tol=1e-10\n"


CASES = [
    ("fp64 relative 256^2 random rectangles", decks.random_rect_deck(3, n=256, steps=3), 3),
    # the one-kernel FP32 iteration (the form graph batches run when nx % 4 == 0)
    ("fp32_ir 256^2 random rectangles",
     decks.random_rect_deck(3, n=256, steps=3, precision="fp32_ir", extra="fp32_kernels=fused\n"), 3),
    ("fp64 reference rule 128^2 reference deck", decks.reference_deck(128, 2, max_iters=700), 2),
    ("fp32_ir 131x77 reference deck",
     decks.reference_deck(131, 2, extra=REL + "precision=fp32_ir\n").replace("y_cells=131", "y_cells=77"), 2),
    ("fp64 relative 67x3 halo 4",
     decks.reference_deck(67, 2, extra=REL).replace("y_cells=67", "y_cells=3").replace("halo_depth=2",
                                                                                       "halo_depth=4"), 2),
]


@pytest.mark.parametrize("name,text,steps", CASES, ids=[c[0].replace(" ", "-") for c in CASES])
def test_graph_batches_bit_identical_to_stream(name, text, steps):
    print(f"case: graph vs stream, {name}, {steps} steps", flush=True)
    ref = _run(text, T.TL_LAUNCH_STREAM, steps)
    got = _run(text, T.TL_LAUNCH_GRAPH, steps)
    assert all(st["iterations"] > 5 for _, st, _ in ref)  # every solve reaches the replays
    _same(got, ref)


def test_graph_recaptured_when_coefficients_change():
    """A new timestep size changes the FP32 face scales held by the captured
    launches: the cached graph must not be replayed with the old ones."""
    print("case: graph vs stream, fp32_ir 192^2 with dt changing per step", flush=True)
    text = decks.random_rect_deck(9, n=192, steps=3, precision="fp32_ir", extra="fp32_kernels=fused\n")
    dts = [0.004, 0.01, 0.004]
    _same(_run(text, T.TL_LAUNCH_GRAPH, 3, dts), _run(text, T.TL_LAUNCH_STREAM, 3, dts))


def test_launch_mode_validated():
    print("case: invalid launch mode", flush=True)
    s = T.Settings.from_text(decks.reference_deck(16, 1)).set(launch=2)
    c = T.initialiseapp(s)
    with pytest.raises(T.TeaLeafError):
        T.step(c, s)


def _golden(name):
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name)) as f:
        return json.load(f)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_graph_mode_c1_deck_matches_golden():
    """The 512^2 reference deck, 10 CG steps with hipGraph batches, against
    the oracle's fixture (not against stream launches): field summaries
    within 1e-10 every step, iterations within max(3, 2%) of the oracle's."""
    print("case: graph mode, golden 512^2 reference deck, 10 CG steps", flush=True)
    g = _golden("c1_512_cg.json")
    s = T.Settings.from_text(g["deck"]).set(launch=T.TL_LAUNCH_GRAPH)
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, 10)
    for t in range(10):
        want = g["intended"]["summary"][t]
        assert _rel(summ[t], want) < 1e-10, (t, summ[t], want)
        wi = g["intended"]["stats"][t]["iterations"]
        assert abs(st[t]["iterations"] - wi) <= max(3, wi // 50), (t, st[t]["iterations"], wi)
        assert st[t]["iterations"] > 64  # the solve ran replays, not only the first stream batch


@pytest.mark.parametrize("precision", ["fp32_ir", "fp64"])
def test_graph_mode_random_rectangles_match_fp64_fixture(precision):
    """A 256^2 random-rectangle deck, 3 steps with hipGraph batches in each
    precision, against the oracle's FP64 CG solved to a relative residual
    of 1e-12 (rect_256_fp64.json): summaries within 1e-9 every step and the
    FP64 residual of every returned solution within 1e-10."""
    print(f"case: graph mode {precision}, 256^2 random rectangles vs the FP64 fixture", flush=True)
    g = _golden("rect_256_fp64.json")
    text = g["deck"].replace("precision=fp32_ir", f"precision={precision}")
    s = T.Settings.from_text(text).set(launch=T.TL_LAUNCH_GRAPH)
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, 3)
    for t in range(3):
        assert st[t]["converged"] == 1 and st[t]["rel_resid"] <= 1e-10, (t, st[t])
        assert st[t]["iterations"] > 64
        assert _rel(summ[t], g["summary"][t]) < 1e-9, (t, summ[t], g["summary"][t])
