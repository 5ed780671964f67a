"""Train the warm-start U-Net on the solver's own trajectories, then measure
what it gains.

    python -m tealeaf_hip.train_warmstart --res 1024 --problems 200 --steps 50 --crops 1 \\
        --train-steps 3000 --out weights/unet_warmstart.npy \\
        --eval-res 4096 --eval-seeds 1000 --eval-steps 50

(the defaults: 200 problems x 50 timesteps at 1024^2, one 128x128 crop per
solve, i.e. 10,000 samples).  With --out, the weights, the loss curve and a
JSON record of the arguments, the dataset size, the final loss and the
evaluation rows are written next to each other.

1. Trajectories: seeded random-rectangle problems (the benchmark family:
   background rho=100, e=1e-4 plus 32 rectangles; dt 0.004) at --res, solved
   to ||u0 - A u|| <= 1e-10 ||u0|| by CG without a warm start; after every
   solve the library records random 128x128 crops of (network input,
   (u_new - u_prev)/s_u) in device memory.
2. Training: Adam on minibatches of crops, all on the device.
3. Evaluation on unseen seeds at --eval-res: the same timesteps solved from
   the previous-step guess and from the learned guess (with its fallback),
   for FP64 CG and for FP32-IR CG; iterations, initial residuals, the start
   chosen (0 previous-step, 1 learned) and solve times per step are printed
   as JSON lines.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import Settings, initialiseapp, step, fieldsummary
from . import decks, unet


def log(**kw):
    print(json.dumps(kw), flush=True)


def generate_and_train(a):
    first = Settings.from_text(decks.random_rect_deck(a.seed0, n=a.res, steps=a.steps, precision="fp32_ir"))
    chunk = initialiseapp(first)
    unet.record(chunk, capacity=a.problems * a.steps * a.crops, crops_per_step=a.crops)
    t0 = time.time()
    for k in range(a.problems):
        s = Settings.from_text(decks.random_rect_deck(a.seed0 + k, n=a.res, steps=a.steps, precision="fp32_ir"))
        if k:
            from . import _lib
            d = s._deck
            _lib.check(_lib.load().tl_initialise(chunk._ctx, d.states, d.nstates), chunk._ctx, "tl_initialise")
        its = []
        for t in range(a.steps):
            st = step(chunk, s)
            its.append(st["iterations"])
        log(phase="trajectories", problem=a.seed0 + k, res=a.res, iterations=its, samples=unet.dataset_size(chunk),
            elapsed_s=round(time.time() - t0, 1))
    t1 = time.time()
    losses = []
    lr = a.lr
    done = 0
    while done < a.train_steps:
        n = min(500, a.train_steps - done)
        if done >= 0.6 * a.train_steps:
            lr = a.lr * 0.3
        if done >= 0.85 * a.train_steps:
            lr = a.lr * 0.1
        loss = unet.train(chunk, steps=n, batch=a.batch, lr=lr, seed=done + 1, loss=a.loss)
        losses.extend(loss.tolist())
        done += n
        log(phase="train", step=done, lr=lr, loss_first=float(np.mean(losses[:20])),
            loss_recent=float(np.mean(loss[-50:])), elapsed_s=round(time.time() - t1, 1))
    w = unet.get_weights(chunk)
    chunk.close()
    return w, losses


def evaluate(w, a):
    rows = []
    for seed in a.eval_seeds:
        for prec in ("fp64", "fp32_ir"):
            res = {}
            for arm in ("none", "learned"):
                warm = arm
                s = Settings.from_text(decks.random_rect_deck(seed, n=a.eval_res, steps=a.eval_steps, precision=prec,
                                                              warm_start=warm))
                c = initialiseapp(s)
                if arm == "learned":
                    unet.set_weights(c, w)
                its, secs, init_r, fb, summ, starts = [], [], [], [], [], []
                for t in range(a.eval_steps):
                    t0 = time.time()
                    st = step(c, s)
                    summ.append(fieldsummary(c))
                    secs.append(time.time() - t0)
                    its.append(st["iterations"])
                    init_r.append(st["init_rel_resid"])
                    fb.append(st["fallback"])
                    starts.append(st["start"])
                res[arm] = dict(iterations=its, seconds=[round(x, 3) for x in secs], init_rel_resid=init_r,
                                 fallback=fb, start=starts, summary=np.array(summ))
                c.close()
            d = np.max(np.abs(res["learned"]["summary"] - res["none"]["summary"]) / np.abs(res["none"]["summary"]))
            row = dict(phase="eval", seed=seed, res=a.eval_res, precision=prec,
                       iterations_prev_guess=res["none"]["iterations"],
                       iterations_learned=res["learned"]["iterations"],
                       seconds_prev_guess=res["none"]["seconds"], seconds_learned=res["learned"]["seconds"],
                       init_rel_resid_prev_guess=res["none"]["init_rel_resid"],
                       init_rel_resid_learned=res["learned"]["init_rel_resid"],
                       fallback=res["learned"]["fallback"], start_learned=res["learned"]["start"],
                       summary_max_rel_diff=float(d))
            log(**row)
            rows.append(row)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--problems", type=int, default=200)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--crops", type=int, default=1)
    ap.add_argument("--seed0", type=int, default=0)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--loss", default="energy", choices=["energy", "residual", "mse"])
    ap.add_argument("--out", default="")
    ap.add_argument("--weights", default="", help="skip training and evaluate these weights")
    ap.add_argument("--eval-res", type=int, default=4096)
    ap.add_argument("--eval-seeds", type=int, nargs="*", default=[1000])
    ap.add_argument("--eval-steps", type=int, default=50)
    a = ap.parse_args(argv)
    losses = None
    if a.weights:
        w = np.load(a.weights, allow_pickle=False)
    else:
        w, losses = generate_and_train(a)
        if a.out:
            os.makedirs(os.path.dirname(a.out) or ".", exist_ok=True)
            np.save(a.out, w, allow_pickle=False)
            np.save(os.path.splitext(a.out)[0] + "_loss.npy", np.asarray(losses, dtype=np.float32),
                    allow_pickle=False)
    rows = evaluate(w, a) if a.eval_seeds else []
    if a.out and losses is not None:
        rec = dict(args=vars(a), samples=a.problems * a.steps * a.crops,
                   loss_first=float(np.mean(losses[:20])), loss_final=float(np.mean(losses[-50:])),
                   eval=[{k: v for k, v in r.items() if k != "summary"} for r in rows])
        with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
            json.dump(rec, f, indent=1)


if __name__ == "__main__":
    sys.exit(main())
