"""Training of the learned block preconditioner of the AMEn local GMRES.

Training pairs come from plain AMEn runs: with collection active, every
block-Jacobi 'c' solve inside the local GMRES records sampled blocks as
(features, normalised exact block correction) on the device.  Runs cover a
set of seeded coefficient draws of the variable-coefficient diffusion
operator; the network is then trained on the GPU by the library's own
forward, backward and Adam kernels (ttf_mlp_train).

    python tt-fort_amd/learned.py [--draws 60] [--per-draw 2000] [--epochs 40] [--rmax 64]

writes the parameters to models/mlp_precond_n64.npy next to this file.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import mpo  # noqa: E402
import ttfort  # noqa: E402

MODEL_PATH = os.path.join(HERE, "models", "mlp_precond_n64.npy")


def collect_pairs(draws, per_draw, d=6, n=64, seed0=1000, rmax=16, tol=1e-6, nswp=3, per_apply=8, verb=0):
    """Run plain AMEn (block-Jacobi 'c') on `draws` seeded coefficient draws
    of the diffusion operator, collecting up to per_draw training pairs from
    each.  Returns host arrays X (256 x N), T (64 x N) and run statistics."""
    Xs, Ts, counts = [], [], []
    t0 = time.perf_counter()
    for k in range(draws):
        crA, ra, nn, _ = mpo.diffusion_7pt(d, n, seed=seed0 + k, coef_rank=3)
        crY, ry = mpo.ones_tt(d, n)
        crX, rx = mpo.random_tt(d, n, 2, seed=seed0 + k)
        ttfort.mlp_collect_begin(per_draw, per_apply)
        try:
            ttfort.tt_amen_solve(nn, ra, crA, ry, crY, rx, crX, tol, kickrank=2, nswp=nswp, prec="c",
                                 nrestart=40, niters=2, rmax=rmax)
        finally:
            ttfort.mlp_collect_end()
        X, T, c = ttfort.mlp_collected(host=True)
        Xs.append(X[:, :c])
        Ts.append(T[:, :c])
        counts.append(c)
        if verb:
            print("draw %d: %d pairs" % (k, c), file=sys.stderr, flush=True)
    X = np.concatenate(Xs, axis=1) if Xs else np.zeros((256, 0), np.float32)
    T = np.concatenate(Ts, axis=1) if Ts else np.zeros((64, 0), np.float32)
    keep = np.all(np.isfinite(X), axis=0) & np.all(np.isfinite(T), axis=0)
    return (np.asfortranarray(X[:, keep]), np.asfortranarray(T[:, keep]),
            dict(draws=draws, pairs=int(keep.sum()), seconds=time.perf_counter() - t0, per_draw=counts))


def train(X, T, epochs=40, batch=1024, lr=1e-3, seed=0):
    """Initialise and train the network on (X, T); returns (params, loss
    history per epoch, seconds)."""
    ttfort.mlp_init(seed)
    t0 = time.perf_counter()
    hist = ttfort.mlp_train(X, T, epochs, batch=batch, lr=lr, seed=seed)
    return ttfort.mlp_get_params(), hist, time.perf_counter() - t0


def save(theta, path=MODEL_PATH):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.save(path, np.asarray(theta, dtype=np.float32), allow_pickle=False)


def load(path=MODEL_PATH):
    """Load saved parameters into the library; returns them (None if the
    file does not exist)."""
    if not os.path.exists(path):
        return None
    theta = np.load(path, allow_pickle=False)
    ttfort.mlp_set_params(theta)
    return theta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--draws", type=int, default=60)
    ap.add_argument("--per-draw", type=int, default=2000)
    ap.add_argument("--rmax", type=int, default=64, help="rank cap of the AMEn runs the pairs come from")
    ap.add_argument("--d", type=int, default=6)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=MODEL_PATH)
    args = ap.parse_args()
    lib = ttfort.load()
    ttfort.check(lib.ttf_init(0), lib)
    X, T, info = collect_pairs(args.draws, args.per_draw, d=args.d, rmax=args.rmax, verb=1)
    print("collected %d pairs from %d draws in %.1f s" % (info["pairs"], args.draws, info["seconds"]), flush=True)
    theta, hist, secs = train(X, T, args.epochs, args.batch, args.lr)
    print("trained %d epochs in %.2f s: loss %s" % (args.epochs, secs, " ".join("%.4g" % h for h in hist)),
          flush=True)
    save(theta, args.out)
    print("saved", args.out)


if __name__ == "__main__":
    main()
