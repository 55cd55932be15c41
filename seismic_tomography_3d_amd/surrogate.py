"""Training and loading of the warm-start network (kernels in csrc/nn.hip and
csrc/surrogate.hip; everything runs on the GPU through the C ABI).

    python -m seismic_tomography_3d_amd.surrogate --out weights/warm_5x5x2.5km.bin

trains on the 201x201x101 / 25 m grid of the batched workload (5 x 5 x 2.5 km):
2000 seeded synthetic models (half layered: 3-7 layers in 1500-4500 m/s, half
linear gradients; each with 1-4 Gaussian anomalies), 4 sources each with
depth U(0, 2 km), solved by the plain FIM of this library.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(PKG, "weights", "warm_5x5x2.5km.bin")


def attach(ctx, case=None, path=None, log=None):
    """Loads trained weights into ctx when the file exists; True if loaded."""
    p = path or WEIGHTS
    if not os.path.exists(p):
        if log:
            log(f"no trained network at {p}")
        return False
    ctx.surrogate_load(p)
    if log:
        log(f"loaded warm-start network {p}")
    return True


def train(ctx, n_models=2000, spm=4, epochs=20, batch=8, lr=2e-3, n_val=256, seed=5,
          nc=32, under_weight=1.0, log=print):
    """Generates the training set on the device and runs Adam epoch by epoch
    with a cosine learning-rate decay (lr -> lr/20).  under_weight > 1 weights
    under-predictions in the loss, training towards an upper bound."""
    ctx.surrogate_init(nc, seed=seed)
    ctx.surrogate_loss(under_weight)
    t0 = time.time()
    ctx.surrogate_make_data(n_models, shots_per_model=spm, zmax_src=2000.0, seed=seed)
    t1 = time.time()
    log(f"training data: {n_models * spm} samples in {t1 - t0:.1f}s")
    hist = np.zeros((epochs, 2), np.float32)
    for e in range(epochs):
        lr_e = lr * (0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * e / max(1, epochs))))
        hist[e] = ctx.surrogate_train(1, batch=batch, lr=float(lr_e), n_val=n_val,
                                      seed=seed + 1 + e)[0]
        log(f"epoch {e}: lr {lr_e:.2e} train {hist[e, 0]:.5f} val {hist[e, 1]:.5f} "
            f"({time.time() - t1:.0f}s)")
    return hist


def main():
    from . import Context, synth
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--models", type=int, default=2000)
    ap.add_argument("--spm", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--history", default=None)
    ap.add_argument("--under-weight", type=float, default=1.0,
                    help="loss weight of under-predictions (> 1 trains towards an upper bound)")
    a = ap.parse_args()
    case = synth.batched_case(1)
    ctx = Context(0)
    ctx.set_model(1.0 / case["vel"], case["h"])
    hist = train(ctx, n_models=a.models, spm=a.spm, epochs=a.epochs, batch=a.batch, lr=a.lr,
                 under_weight=a.under_weight, log=lambda m: print(m, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    ctx.surrogate_save(a.out)
    print(f"saved {a.out}")
    if a.history:
        with open(a.history, "w") as f:
            json.dump({"train": hist[:, 0].tolist(), "val": hist[:, 1].tolist()}, f)


if __name__ == "__main__":
    main()
