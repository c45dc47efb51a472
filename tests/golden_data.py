"""Loaders of the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from numpy einsum and the dense brute force)."""
import os

import numpy as np

from qutree_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def node_ops():
    return np.load(os.path.join(GOLDEN, "node_ops.npz"), allow_pickle=False)


def sweep8():
    """(tree text, Hamiltonian, fixture arrays) of the 8-mode golden sweep."""
    d = np.load(os.path.join(GOLDEN, "sweep8.npz"), allow_pickle=False)
    H = W.Hamiltonian(coeff=[complex(c) for c in d["coeff"]], term_off=[int(x) for x in d["term_off"]],
                      fmode=[int(x) for x in d["fmode"]], fkind=[int(x) for x in d["fkind"]],
                      fpool=[int(x) for x in d["fpool"]], pool=[float(x) for x in d["pool"]])
    return str(d["tree_text"]), H, d
