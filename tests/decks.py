"""Small seeded decks shared by the tests (sizes the oracle finishes in
seconds).  Everything is generated; nothing reads the reference tree."""
import numpy as np

from brownian_dynamics_cuda_amd import fiber as Fb


def deck(kind, hi_all=0):
    """kind: 'tiny' (3 cores, tails, LH), 'small' (6 cores, tails, LH),
    'cfg1' (12 cores, tails, no LH, the 12-nucleosome recipe)."""
    if kind == "tiny":
        return Fb.make_fiber(3, seed=11, with_lh=True, hi_all=hi_all)
    if kind == "small":
        return Fb.make_fiber(6, seed=12, with_lh=True, hi_all=hi_all)
    if kind == "cfg1":
        return Fb.make_config(1, hi_all=hi_all)
    raise KeyError(kind)


def rel_err(a, b):
    """max |a - b| / max |b| (norm-wise relative error)."""
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    if b.size == 0:
        return 0.0
    den = np.max(np.abs(b))
    if den == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / den)
