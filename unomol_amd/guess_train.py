"""Training of the learned initial-density network on densities this engine converges.

Pipeline: seeded water clusters (the project's generator) -> GPU SCF (FP64, DIIS) ->
atom-pair descriptors and density blocks in the pair frames -> network trained on the
GPU (BF16 matrix-core forward/backward, Adam) -> used by findEnergy(guess=2).

usage: python -m unomol_amd.guess_train OUT.bin [--clusters N] [--min-w A] [--max-w B]
"""
import argparse
import os
import sys
import tempfile
import time

import numpy as np

from . import api


def converged_density(deck_path, eps=1e-9):
    b = api.Basis(deck_path)
    t = api.TwoElectronInts(b)
    hf = api.RestrictedHartreeFock(b, t)
    hf.findEnergy(accel=2, eps=eps, cflag=3, maxits=100)
    return b, hf


def make_training_set(n_clusters=200, min_w=2, max_w=12, seed=1000, basis="6-31G*", log=None):
    """Descriptors, blocks and masks from converged densities of n_clusters clusters with
    min_w..max_w waters (seeds seed, seed+1, ...)."""
    Xs, Ys, Ms = [], [], []
    rng = np.random.default_rng(seed)
    tmp = tempfile.mkdtemp()
    t0 = time.time()
    for k in range(n_clusters):
        nw = int(rng.integers(min_w, max_w + 1))
        path = os.path.join(tmp, "c%d.dat" % k)
        api.generate_deck(path, "water_cluster", nw, seed + k, basis, 0, 1e-9, 100)
        b, hf = converged_density(path)
        X, Y, M = api.guess_samples(b.deck, hf.P)
        Xs.append(X)
        Ys.append(Y)
        Ms.append(M)
        if log and (k + 1) % 20 == 0:
            log("training set: %d/%d clusters, %d samples, %.1f s" % (k + 1, n_clusters, sum(len(x) for x in Xs),
                                                                      time.time() - t0))
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)


def train_guess(X, Y, M, steps=3000, batch=1024, lr=2e-3, seed=0, width=128, hidden=3):
    net = api.MLP(api.GUESS_IN, width, hidden, api.GUESS_OUT, seed=seed)
    loss = net.train(X, Y, M, steps=steps, batch=batch, lr=lr, seed=seed)
    return net, loss


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--clusters", type=int, default=200)
    ap.add_argument("--min-w", type=int, default=2)
    ap.add_argument("--max-w", type=int, default=12)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=1000)
    args = ap.parse_args()

    def log(m):
        print("[guess-train] " + m, flush=True)

    X, Y, M = make_training_set(args.clusters, args.min_w, args.max_w, args.seed, log=log)
    log("samples %d" % len(X))
    t0 = time.time()
    net, loss = train_guess(X, Y, M, steps=args.steps)
    log("trained %d steps in %.1f s, masked MSE %.3e" % (args.steps, time.time() - t0, loss))
    net.save(args.out)
    log("saved " + args.out)


if __name__ == "__main__":
    sys.exit(main())
