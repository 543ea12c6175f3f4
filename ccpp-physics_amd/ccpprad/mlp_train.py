"""Train the LW gas-optics emulator on the GPU and write its weights.

    python -m ccpprad.mlp_train --samples 4000000 --epochs 24 --out weights.npz

Samples: 4M layers from seeded profiles drawn like the all-sky LW workload.
Targets come from the k-distribution gas optics (GPU table path). Training
uses the library's BF16 MFMA forward/backward and Adam. The script prints the
held-out emulation error of ln(tau/col_dry) and of the Planck fractions.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=4_000_000)
    ap.add_argument("--epochs", type=int, default=24)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=20261204)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    from . import mlp, synthetic as syn

    t0 = time.time()
    kd = syn.make_kdist(True)
    X, Y = mlp.make_training_set(kd, nsamples=a.samples, seed=a.seed, torch=torch)
    print(f"# samples {X.shape[0]} generated in {time.time() - t0:.1f} s", flush=True)
    nval = min(200_000, X.shape[0] // 20)
    Xv, Yv = X[-nval:], Y[-nval:]
    Xt, Yt = X[:-nval], Y[:-nval]
    e = mlp.LWEmulator(seed=a.seed)
    e.set_norm(**mlp.fit_norm(Xt, Yt))
    t1 = time.time()
    hist = mlp.train(e, Xt, Yt, torch, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed,
                     log=lambda s: print("#", s, flush=True))
    torch.cuda.synchronize()
    ttrain = time.time() - t1
    # held-out error (standardised forward of the training path)
    n = e.norm
    Xs = ((Xv - torch.from_numpy(n["xmean"]).cuda().double()) / torch.from_numpy(n["xstd"]).cuda().double()).float()
    Yp = torch.zeros((Xs.shape[0], 256), dtype=torch.float32, device="cuda")
    e.forward(Xs.contiguous(), Yp)
    y = Yp * torch.from_numpy(n["ystd"]).cuda() + torch.from_numpy(n["ymean"]).cuda()
    dl = (y[:, :128] - Yv[:, :128]).abs()
    dp = (y[:, 128:] - Yv[:, 128:]).abs()
    rep = {"train_seconds": ttrain, "epochs": a.epochs, "samples": int(Xt.shape[0]), "loss_history": hist,
           "val_ln_tau_mae": float(dl.mean()), "val_ln_tau_p99": float(torch.quantile(dl.flatten()[:1_000_000], 0.99)),
           "val_pfrac_mae": float(dp.mean())}
    print(json.dumps(rep), flush=True)
    out = a.out or mlp.DEFAULT_WEIGHTS
    os.makedirs(os.path.dirname(out), exist_ok=True)
    e.save(out)
    print("# wrote", out, flush=True)


if __name__ == "__main__":
    sys.exit(main())
