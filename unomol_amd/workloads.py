"""The synthetic workloads of the benchmarks and tests, built with the project's own deck
generator (uno_deck_generate): seeded water clusters and a pool of small molecules."""
import os

import numpy as np

from . import api


def batched_pool(directory, count=512, seed=5, bases=("6-31G", "6-31G*"), eps=1e-8, maxits=200):
    """Deck paths of the batched-SCF pool: seeded distortions (bonds +-5 %, angles +-5 deg)
    of H2, HF, H2O, NH3, CH4, N2, CO, CO2 and C2H2 and 1-8-water clusters, each in one of
    `bases` (drawn per molecule), every other one a +1 cation (UHF doublet), the rest
    neutral (RHF).  Sizes run from 4 to 152 Cartesian functions."""
    rng = np.random.default_rng(seed)
    paths = []
    for i in range(count):
        charge = i % 2
        basis = bases[int(rng.integers(0, len(bases)))]
        path = os.path.join(os.fspath(directory), "m%03d.dat" % i)
        if rng.uniform() < 0.5:
            api.generate_deck(path, "molecule", int(rng.integers(0, 9)), seed * 10000 + i, basis, charge, eps, maxits)
        else:
            api.generate_deck(path, "water_cluster", int(rng.integers(1, 9)), seed * 10000 + i, basis, charge, eps,
                              maxits)
        paths.append(path)
    return paths


def water_cluster(path, n, seed, basis="6-31G*", charge=0, eps=1e-8, maxits=200):
    """One seeded water cluster deck (n waters at 1 g/cm^3, O-O >= 2.5 A, random rotations)."""
    api.generate_deck(path, "water_cluster" if n > 1 else "water", n, seed, basis, charge, eps, maxits)
    return path
