"""Training of the learned initial-density network on densities this engine converges.

Pipeline: seeded water clusters (the project's generator) -> GPU SCF (FP64, DIIS) ->
atom-pair descriptors and density blocks in the pair frames -> network trained on the
GPU (BF16 matrix-core forward/backward, Adam) -> used by findEnergy(guess=2).

usage: python -m unomol_amd.guess_train OUT.bin [--clusters N] [--min-w A] [--max-w B]
       python -m unomol_amd.guess_train OUT.bin --batched [--small N] [--large M] [--cap K]
           (N clusters of 2-13 waters through the batched SCF, M clusters of 14-32 waters,
           at most K pair samples per cluster)

The shipped network models/guess_631gs.bin (6-31G*) was made by
    python -m unomol_amd.guess_train unomol_amd/models/guess_631gs.bin --batched \
        --small 4000 --large 1000 --seed 20000 --cap 300 --steps 20000 --batch 2048 --lr 2e-3
(1.03 million pair samples from 5000 clusters; the 64-water clusters it is tested on are
generated from other seeds).
"""
import argparse
import os
import sys
import tempfile
import time

import numpy as np

from . import api


def converged_density(deck_path, eps=1e-9, guess_net=None):
    """FP64 SCF (DIIS, full builds) converged to |dE| < eps; with guess_net the SCF starts
    from that network's guess (only the start changes, not the converged density)."""
    b = api.Basis(deck_path)
    t = api.TwoElectronInts(b)
    opts = dict(accel=2, eps=eps, cflag=3, maxits=100)
    if guess_net is not None:
        api.set_guess_model(t, guess_net)
        opts["guess"] = 2
    hf = api.RestrictedHartreeFock(b, t)
    hf.findEnergy(**opts)
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


def _keep(rng, X, Y, M, cap):
    if cap and len(X) > cap:
        idx = np.sort(rng.choice(len(X), cap, replace=False))
        return X[idx], Y[idx], M[idx]
    return X, Y, M


def make_training_set_batched(n_small=4000, n_large=1000, small_w=(2, 13), large_w=(14, 32), seed=20000,
                              basis="6-31G*", cap=300, batch=512, log=None, n_xl=0, xl_w=(40, 64),
                              guess_net=None):
    """A larger training set from the batched SCF: n_small clusters of small_w waters run
    `batch` at a time through one batched SCF (FP64, DIIS, density and energy converged to
    1e-9), n_large clusters of large_w waters one by one; each cluster's atom pairs are
    sub-sampled to at most `cap` samples (uniformly), so that many clusters contribute.
    n_xl clusters of xl_w waters (bulk-like environments) follow one by one, their SCF
    started from guess_net when given.  Seeds seed, seed+1, ... (the held-out 64-water
    clusters use other seeds)."""
    rng = np.random.default_rng(seed)
    tmp = tempfile.mkdtemp()
    Xs, Ys, Ms = [], [], []
    t0 = time.time()
    k = 0
    # small clusters, batched
    for start in range(0, n_small, batch):
        paths = []
        for j in range(min(batch, n_small - start)):
            nw = int(rng.integers(small_w[0], small_w[1] + 1))
            path = os.path.join(tmp, "s%d.dat" % k)
            api.generate_deck(path, "water_cluster", nw, seed + k, basis, 0, 1e-9, 150)
            paths.append(path)
            k += 1
        b = api.Batch(paths)
        e, res = b.run(maxits=150, eps=1e-9, cflag=2, accel=2)
        for i, path in enumerate(paths):
            if not res[i].converged:
                continue
            d = b.decks[i]
            n = d["norb"]
            P = b.density(i)[np.tril_indices(n)]
            X, Y, M = _keep(rng, *api.guess_samples(d, np.ascontiguousarray(P)), cap)
            Xs.append(X)
            Ys.append(Y)
            Ms.append(M)
        del b
        if log:
            log("training set: %d/%d small clusters, %d samples, %.1f s" %
                (k, n_small, sum(len(x) for x in Xs), time.time() - t0))
    # larger clusters, one SCF each
    for j in range(n_large):
        nw = int(rng.integers(large_w[0], large_w[1] + 1))
        path = os.path.join(tmp, "l%d.dat" % k)
        api.generate_deck(path, "water_cluster", nw, seed + k, basis, 0, 1e-9, 150)
        k += 1
        bb, hf = converged_density(path)
        if not hf.result.converged:
            continue
        X, Y, M = _keep(rng, *api.guess_samples(bb.deck, hf.P), cap)
        Xs.append(X)
        Ys.append(Y)
        Ms.append(M)
        if log and (j + 1) % 100 == 0:
            log("training set: %d/%d larger clusters, %d samples, %.1f s" %
                (j + 1, n_large, sum(len(x) for x in Xs), time.time() - t0))
    for j in range(n_xl):
        nw = int(rng.integers(xl_w[0], xl_w[1] + 1))
        path = os.path.join(tmp, "x%d.dat" % k)
        api.generate_deck(path, "water_cluster", nw, seed + k, basis, 0, 1e-9, 150)
        k += 1
        bb, hf = converged_density(path, guess_net=guess_net)
        if not hf.result.converged:
            continue
        X, Y, M = _keep(rng, *api.guess_samples(bb.deck, hf.P), cap)
        Xs.append(X)
        Ys.append(Y)
        Ms.append(M)
        if log and (j + 1) % 20 == 0:
            log("training set: %d/%d cluster-scale clusters, %d samples, %.1f s" %
                (j + 1, n_xl, sum(len(x) for x in Xs), time.time() - t0))
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
    ap.add_argument("--batched", action="store_true")
    ap.add_argument("--small", type=int, default=4000)
    ap.add_argument("--large", type=int, default=1000)
    ap.add_argument("--xl", type=int, default=0, help="clusters of 40-64 waters")
    ap.add_argument("--xl-guess", default="", help="network starting the SCFs of those clusters")
    ap.add_argument("--cap", type=int, default=300)
    ap.add_argument("--batch", type=int, default=1024, help="training minibatch")
    ap.add_argument("--lr", type=float, default=2e-3)
    args = ap.parse_args()

    def log(m):
        print("[guess-train] " + m, flush=True)

    if args.batched:
        gn = api.MLP(path=args.xl_guess) if args.xl_guess else None
        X, Y, M = make_training_set_batched(args.small, args.large, seed=args.seed, cap=args.cap, log=log,
                                            n_xl=args.xl, guess_net=gn)
    else:
        X, Y, M = make_training_set(args.clusters, args.min_w, args.max_w, args.seed, log=log)
    log("samples %d" % len(X))
    t0 = time.time()
    net, loss = train_guess(X, Y, M, steps=args.steps, batch=args.batch, lr=args.lr)
    log("trained %d steps in %.1f s, masked MSE %.3e" % (args.steps, time.time() - t0, loss))
    net.save(args.out)
    log("saved " + args.out)


if __name__ == "__main__":
    sys.exit(main())
