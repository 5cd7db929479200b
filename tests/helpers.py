"""Test inputs: seeded hand-made snapshots that stress the AMR edge cases."""
import numpy as np

from ahf_amd import dumps


def fuzz_snapshot(path, seed, non_pow2=False):
    """Random small snapshot: background + Gaussian clusters, some placed on
    the periodic faces, some particles exactly on cell faces; random
    thresholds and depth.  non_pow2 draws domain grids that are even but not
    powers of two (AHF takes any LgridDomain).  Returns a dict of the
    parameters."""
    rng = np.random.default_rng(seed)
    L = int(rng.choice([6, 10, 12, 20, 24] if non_pow2 else [4, 8, 8, 16, 16, 32]))
    n_bg = int(rng.integers(0, L ** 3 + 1))
    parts = [rng.random((n_bg, 3))]
    for _ in range(int(rng.integers(1, 6))):
        c = rng.random(3)
        if rng.random() < 0.5:
            c[rng.integers(0, 3)] = rng.choice([0.0, 0.999, 0.5 / L, 1 - 0.5 / L])
        n = int(rng.integers(10, 3000))
        s = rng.choice([0.002, 0.01, 0.03, 0.1])
        parts.append((c + s * rng.standard_normal((n, 3))) % 1.0)
    if rng.random() < 0.3:
        k = int(rng.integers(1, 200))
        parts.append(rng.integers(0, 4 * L, (k, 3)) / (4.0 * L))
    pos = np.concatenate(parts).astype(np.float32)
    pos[pos >= 1.0] = 0.0
    rng.shuffle(pos)
    nth = float(rng.choice([1.0, 2.0, 3.0, 5.0, 8.0]))
    Lmax = L << int(rng.integers(1, 7))
    nref = float(rng.choice([nth, 2.5, 5.0]))
    dumps.write_snapshot(path, pos, L_dom=L, L_max=Lmax, Nth_dom=nth, Nth_ref=nref)
    return dict(L=L, n=len(pos), Nth_dom=nth, Nth_ref=nref, L_max=Lmax)
