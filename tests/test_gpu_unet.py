"""GPU parity of the warm-start U-Net (BF16 activations and weights, FP32
accumulation on MFMA) against the FP64 NumPy oracle (oracle/unet_ref.py).

Tolerances are those of BF16 storage: inputs and weights are rounded to
BF16 on both sides, so the remaining difference is the rounding of the
activations and gradients between layers (relative ~2^-8 per rounding).
"""
import numpy as np
import pytest

import tealeaf_hip as T
import unet_ref as U
from tealeaf_hip import _lib, decks, unet

pytestmark = pytest.mark.gpu


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def chunk():
    s = T.Settings.from_text(decks.reference_deck(16, 1))
    return T.initialiseapp(s)


def _case(seed, B=2, H=16, W=24):
    rng = np.random.default_rng(seed)
    p = U.init_params(seed).astype(np.float32)
    x = rng.standard_normal((B, H, W, 6)).astype(np.float32)
    x16 = np.zeros((B, H, W, 16))
    x16[..., :6] = bf16(x)
    return p, x, x16, rng


def test_param_count():
    assert unet.param_count() == U.param_count() == 118833


@pytest.mark.parametrize("shape", [(2, 16, 24), (1, 32, 32), (1, 8, 72)])
def test_forward_matches_oracle(chunk, shape):
    print(f"case: unet forward {shape}", flush=True)
    p, x, x16, _ = _case(3, *shape)
    unet.set_weights(chunk, p)
    out = unet.evaluate(chunk, x)
    ref, _ = U.forward(bf16(p), x16)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 2e-2, rel
    np.testing.assert_array_equal(unet.get_weights(chunk), p)


def test_gradients_match_oracle(chunk):
    print("case: unet gradients 2x16x24", flush=True)
    p, x, x16, rng = _case(5)
    unet.set_weights(chunk, p)
    target = rng.standard_normal(x.shape[:3]) * 0.3
    out, loss, g = unet.evaluate(chunk, x, target)
    L, gref = U.loss_and_grads(bf16(p), x16, target)
    assert abs(loss - L) / L < 2e-2
    cos = float(np.dot(g, gref) / (np.linalg.norm(g) * np.linalg.norm(gref)))
    assert cos > 0.999, cos
    assert np.linalg.norm(g - gref) / np.linalg.norm(gref) < 5e-2
    # per layer: the largest gradients agree too
    o = 0
    for ci, _, co, _, _ in U.LAYERS:
        n = 9 * ci * co + co
        a, b = g[o:o + n], gref[o:o + n]
        assert np.linalg.norm(a - b) <= 8e-2 * np.linalg.norm(b) + 1e-6, (ci, co)
        o += n


def test_records_and_trains():
    print("case: unet record 256^2 x 4 steps, train 80 steps", flush=True)
    s = T.Settings.from_text(decks.random_rect_deck(21, n=256, steps=4))
    c = T.initialiseapp(s)
    unet.record(c, capacity=64, crops_per_step=8)
    T.diffuse(c, s, 4)
    assert unet.dataset_size(c) == 32
    loss = unet.train(c, steps=80, batch=8, lr=2e-3)
    assert np.all(np.isfinite(loss))
    assert np.mean(loss[-10:]) < 0.5 * np.mean(loss[:5]), (loss[:5], loss[-10:])


def test_learned_start_falls_back_to_previous_guess():
    """The learned warm start is used only when its residual beats the
    previous-step start's.  An all-zero network predicts no change, i.e. the
    previous-step start itself: not better, so the solve falls back to it and
    runs exactly the solve without a warm start."""
    print("case: learned start selection 192^2, zero network", flush=True)
    text = decks.random_rect_deck(31, n=192, steps=2)
    s0 = T.Settings.from_text(text)
    base = T.initialiseapp(s0)
    sb, stb = T.diffuse(base, s0, 2)
    s1 = T.Settings.from_text(text.replace("warm_start=none", "warm_start=learned"))
    s1.set(warm_policy=_lib.TL_WARM_POLICY_EVERY)
    c = T.initialiseapp(s1)
    unet.set_weights(c, np.zeros(unet.param_count(), dtype=np.float32))
    sl, stl = T.diffuse(c, s1, 2)
    for t in range(2):
        assert stl[t]["fallback"] == 1 and stl[t]["start"] == _lib.TL_START_PREVIOUS, stl[t]
        assert stl[t]["iterations"] == stb[t]["iterations"]
        assert np.array_equal(sl[t], sb[t])
    assert stb[0]["start"] == _lib.TL_START_PREVIOUS and stb[0]["fallback"] == 0


def test_trained_network_gain_on_unseen_problem():
    """What the committed trained network gains over the previous-step start
    on an unseen 256^2 problem over 3 steps.  The selection uses the
    prediction only when its residual is smaller; either way the solve meets
    the tolerance and the answers agree.  The iteration counts of both runs
    are printed: on this problem family the gain is about zero (the
    network's receptive field is far smaller than a step's diffusion
    length)."""
    import os
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tealeaf.jl_amd", "weights",
                        "unet_warmstart.npy")
    if not os.path.exists(path):
        pytest.skip("no trained weights committed")
    print("case: trained network vs previous-step start, 256^2 seed 2001, 3 steps", flush=True)
    text = decks.random_rect_deck(2001, n=256, steps=3)
    s0 = T.Settings.from_text(text)
    base = T.initialiseapp(s0)
    sb, stb = T.diffuse(base, s0, 3)
    s1 = T.Settings.from_text(text.replace("warm_start=none", "warm_start=learned"))
    s1.set(warm_policy=_lib.TL_WARM_POLICY_EVERY)
    c = T.initialiseapp(s1)
    unet.load_weights(c, path)
    sl, stl = T.diffuse(c, s1, 3)
    print("  iterations without / with the network:", [x["iterations"] for x in stb],
          [x["iterations"] for x in stl], "start:", [x["start"] for x in stl], flush=True)
    for t in range(3):
        assert stl[t]["rel_resid"] <= 1e-10 and stb[t]["rel_resid"] <= 1e-10
        np.testing.assert_allclose(sl[t], sb[t], rtol=1e-9)
        if stl[t]["start"] == _lib.TL_START_WARM:
            assert stl[t]["fallback"] == 0 and stl[t]["init_rel_resid"] < stb[t]["init_rel_resid"]
        else:
            assert stl[t]["fallback"] == 1


def test_learned_start_backoff_after_rejections():
    """TL_WARM_POLICY_BACKOFF (the default): the network runs on the first
    solve and after every accepted prediction; after a rejection the next
    attempt waits 2, 4, 8, then 16 solves.  A zero network is always rejected,
    so over 16 steps it runs at solves 1, 3, 7 and 15; the other solves start
    from the previous-step guess without evaluating it, and every solve meets
    the same tolerance as the run without a warm start."""
    print("case: learned start back-off, 64^2 zero network, 16 steps", flush=True)
    text = decks.random_rect_deck(32, n=64, steps=16)
    base = T.initialiseapp(T.Settings.from_text(text))
    sb, stb = T.diffuse(base, T.Settings.from_text(text), 16)
    s1 = T.Settings.from_text(text.replace("warm_start=none", "warm_start=learned"))
    c = T.initialiseapp(s1)
    unet.set_weights(c, np.zeros(unet.param_count(), dtype=np.float32))
    sl, stl = T.diffuse(c, s1, 16)
    attempted = [i + 1 for i, x in enumerate(stl) if x["warm_attempted"]]
    assert attempted == [1, 3, 7, 15], attempted
    for t in range(16):
        assert stl[t]["fallback"] == stl[t]["warm_attempted"] and stl[t]["start"] == _lib.TL_START_PREVIOUS
        assert (stl[t]["warm_ms"] > 0) == bool(stl[t]["warm_attempted"])
        assert stl[t]["iterations"] == stb[t]["iterations"] and np.array_equal(sl[t], sb[t])
