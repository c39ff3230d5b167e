"""Loading helpers for the golden fixtures written by tests/golden/make_golden.py."""

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ("smoke12", "moving20", "spin16", "sparse16", "perm24")


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False)


def move_sparse16(particles, k):
    """The in-place motion make_golden applies before every step of sparse16."""
    for j, q in enumerate(particles):
        q.x += 0.013 * (j + 1)
        q.z -= 0.007 * (j + 1)
