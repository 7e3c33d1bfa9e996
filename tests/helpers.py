"""Seeded test inputs: confined random-walk chains relaxed by the oracle."""
import numpy as np
from scipy.spatial import cKDTree

import pyoracle as po

RCO = float(np.float32(1.154701))
MIS = 0.838732


def radius_for(N, cvf):
    """Nucleus radius for chromatin volume fraction cvf (chrdat.cu:57-58 inverted)."""
    return RCO / 2 + RCO / 2 * (N / cvf) ** (1.0 / 3.0)


def walk(N, R_n, seed, chain_off=None):
    rng = np.random.default_rng(seed)
    chain_off = [0, N] if chain_off is None else list(chain_off)
    starts = set(chain_off[:-1])
    r = np.zeros((N, 3))
    lim = R_n - 1.0
    for i in range(N):
        while True:
            if i in starts:
                v = rng.normal(size=3)
                p = v / np.linalg.norm(v) * lim * rng.random() ** (1 / 3)
            else:
                v = rng.normal(size=3)
                p = r[i - 1] + v / np.linalg.norm(v)
            if np.linalg.norm(p) < lim:
                break
        r[i] = p
    return r


def nonbonded_min_distance(r):
    tree = cKDTree(r)
    pairs = tree.query_pairs(1.5, output_type="ndarray")
    pairs = pairs[np.abs(pairs[:, 0] - pairs[:, 1]) > 1]
    if len(pairs) == 0:
        return np.inf
    return np.linalg.norm(r[pairs[:, 0]] - r[pairs[:, 1]], axis=1).min()


def relaxed(N, R_n, seed, chain_off=None, max_rounds=200):
    """Walk + noise-free soft-repulsion steps (ccp's ICG relaxation) until
    no non-bonded pair is closer than 0.838732."""
    r = walk(N, R_n, seed, chain_off)
    sysi = po.System(N, R_n, chain_off=chain_off, ff=po.FF_ICG, sd=0.0)
    sysi.p.sd = 0.0
    for _ in range(max_rounds):
        if nonbonded_min_distance(r) >= MIS:
            return r
        r = sysi.run(r, 0, 64)
    raise RuntimeError("relaxation did not converge")


def pattern(N, seed, frac_het=0.5, mean=20):
    """Alternating A/B domains of geometric lengths; returns ccp flags."""
    rng = np.random.default_rng(seed)
    fl = np.zeros(N, np.uint8)
    i, het = 0, bool(rng.random() < frac_het)
    while i < N:
        L = int(rng.geometric(1.0 / mean))
        if het:
            fl[i:i + L] = po.HET | po.LAD
        i += L
        het = not het
    return fl


def lbs_sites(R_n, n, seed):
    """n lamina binding sites uniformly on the sphere of radius R_n - rco."""
    d = np.random.default_rng(seed).normal(size=(n, 3))
    return (R_n - po.CCP_CONST["rco"]) * d / np.linalg.norm(d, axis=1, keepdims=True)
