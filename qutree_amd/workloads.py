"""Synthetic workload descriptions: trees in QuTree's compact text format and
sum-of-products Hamiltonians as flat arrays.

The tree text is the format read by ``Tree(istream)`` / ``Node::initialize``
(src/TreeShape/Node.cpp:101-165): an upper node is ``nstates -nchildren``
followed by its children, a leaf is ``dim type mode``; after the tree, one
``omega r0 wfr0 wfomega`` line per mode (src/TreeShape/Tree.cpp:137-152).

A Hamiltonian is a list of product terms (``MultiLeafOperator``,
src/TreeOperators/MultiLeafOperator.cpp) with complex coefficients
(``SumOfProductsOperator::push_back``).  Each factor acts on one mode and is
one of the DVR leaf operators of ``LeafInterface`` (x, x^2, kin, p, identity;
src/TreeShape/LeafTypes/DVRBasis.cpp:8-58), an explicit complex matrix
(``LeafMatrix``, src/TreeOperators/LeafMatrix.cpp:37-41), or a ``LeafMatrix``
assembled on the leaf's own DVR grid from a short polynomial
``a_T T + a_1 x + a_2 x^2 + a_4 x^4`` (FACTOR_POLY).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

FACTOR_X, FACTOR_X2, FACTOR_KIN, FACTOR_P, FACTOR_ID, FACTOR_MATRIX, FACTOR_POLY = range(7)


@dataclass
class Hamiltonian:
    """Flat SOP description: term l has coefficient coeff[l] and factors
    term_off[l]:term_off[l+1] of (mode, kind, pool offset)."""

    coeff: List[complex] = field(default_factory=list)
    term_off: List[int] = field(default_factory=lambda: [0])
    fmode: List[int] = field(default_factory=list)
    fkind: List[int] = field(default_factory=list)
    fpool: List[int] = field(default_factory=list)
    pool: List[float] = field(default_factory=list)

    def add_term(self, c: complex, factors: Sequence[tuple]) -> None:
        """factors: (mode, kind) or (mode, FACTOR_POLY, (aT, a1, a2, a4)) or
        (mode, FACTOR_MATRIX, complex ndarray d x d)."""
        for f in factors:
            mode, kind = int(f[0]), int(f[1])
            off = 0
            if kind == FACTOR_POLY:
                off = len(self.pool) // 2
                a = [float(x) for x in f[2]]
                # stored as 2 complex slots = 4 doubles
                self.pool.extend(a)
            elif kind == FACTOR_MATRIX:
                m = np.asarray(f[2], dtype=np.complex128)
                off = len(self.pool) // 2
                flat = np.asfortranarray(m).ravel(order="F")
                inter = np.empty(2 * flat.size)
                inter[0::2] = flat.real
                inter[1::2] = flat.imag
                self.pool.extend(inter.tolist())
            self.fmode.append(mode)
            self.fkind.append(kind)
            self.fpool.append(off)
        self.coeff.append(complex(c))
        self.term_off.append(len(self.fmode))

    @property
    def nterms(self) -> int:
        return len(self.coeff)

    def arrays(self):
        """(coeff interleaved f64, term_off i64, fmode i64, fkind i64, fpool i64, pool f64)."""
        c = np.empty(2 * self.nterms)
        c[0::2] = [z.real for z in self.coeff]
        c[1::2] = [z.imag for z in self.coeff]
        pool = np.asarray(self.pool if self.pool else [0.0, 0.0], dtype=np.float64)
        return (c, np.asarray(self.term_off, dtype=np.int64), np.asarray(self.fmode, dtype=np.int64),
                np.asarray(self.fkind, dtype=np.int64), np.asarray(self.fpool, dtype=np.int64), pool)


def balanced_tree_text(n_leaves: int, dim_leaf: int, spf_by_layer: Sequence[int],
                       omega=1.0, r0: float = 0.0, wfr0: float = 0.0,
                       wfomega=1.0, nstates: int = 1) -> str:
    """Balanced binary tree over ``n_leaves`` (a power of two) HO-DVR leaves.

    spf_by_layer[0] is the SPF count of the bottom nodes, spf_by_layer[k] that
    of nodes k layers above them; the top node carries ``nstates`` states.
    ``omega`` and ``wfomega`` are one value for every mode or one per mode.
    Leaf modes run 0..f-1 from left to right, as ``Tree::resetLeafModes``
    numbers them (src/TreeShape/Tree.cpp:35-46)."""
    if n_leaves < 2 or n_leaves & (n_leaves - 1):
        raise ValueError("n_leaves must be a power of two >= 2")
    nlayers = int(np.log2(n_leaves))  # upper layers above the bottom nodes
    if len(spf_by_layer) < nlayers:
        raise ValueError(f"need {nlayers} SPF counts (bottom + {nlayers - 1} intermediate layers)")
    lines: List[str] = []
    mode = [0]

    def node(height: int, indent: int) -> None:
        # height 0 = bottom node
        tab = "\t" * indent
        if height == nlayers:
            lines.append(f"{tab}{nstates}\t-2")
        else:
            lines.append(f"{tab}{spf_by_layer[height]}\t-{1 if height == 0 else 2}")
        if height == 0:
            lines.append(f"{tab}\t{dim_leaf}\t0\t{mode[0]}")
            mode[0] += 1
            return
        node(height - 1, indent + 1)
        node(height - 1, indent + 1)

    node(nlayers, 0)
    om = np.broadcast_to(np.asarray(omega, dtype=float), (n_leaves,))
    wo = np.broadcast_to(np.asarray(wfomega, dtype=float), (n_leaves,))
    for k in range(n_leaves):
        lines.append(f"{float(om[k])!r} {r0} {wfr0} {float(wo[k])!r}")
    return "\n".join(lines) + "\n"


def _pairs(f: int):
    return [(i, j) for i in range(f) for j in range(i + 1, f)]


def coupled_ho(f: int, seed: int, omega_range=(0.8, 1.2), c_sigma=0.02, one_mode_matrix=False) -> Hamiltonian:
    """H = sum_i T_i + 1/2 w_i^2 q_i^2 + sum_{i<j} c_ij q_i q_j with
    w_i ~ U[omega_range], c_ij ~ N(0, c_sigma^2) drawn from a seeded
    numpy PCG64 generator.  Built through the operator API like CoupledHO
    (contrib/Hamiltonians/Operators/CoupledHO/CoupledHO.cpp:7-66): either as
    separate kin and x^2 terms, or as one LeafMatrix T + 1/2 w^2 x^2 per mode."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(omega_range[0], omega_range[1], size=f)
    pairs = _pairs(f)
    c = rng.normal(0.0, c_sigma, size=len(pairs))
    H = Hamiltonian()
    if one_mode_matrix:
        for i in range(f):
            H.add_term(1.0, [(i, FACTOR_POLY, (1.0, 0.0, 0.5 * w[i] ** 2, 0.0))])
    else:
        for i in range(f):
            H.add_term(1.0, [(i, FACTOR_KIN)])
        for i in range(f):
            H.add_term(0.5 * w[i] ** 2, [(i, FACTOR_X2)])
    for (i, j), cij in zip(pairs, c):
        H.add_term(cij, [(i, FACTOR_X), (j, FACTOR_X)])
    return H


def coupled_quartic(f: int, seed: int, quartic=0.05, c_sigma=0.01) -> Hamiltonian:
    """H = sum_i (T_i + 1/2 q_i^2 + quartic q_i^4) + sum_{i<j} c_ij q_i q_j,
    c_ij ~ N(0, c_sigma^2); one LeafMatrix per mode plus the pair terms."""
    rng = np.random.default_rng(seed)
    pairs = _pairs(f)
    c = rng.normal(0.0, c_sigma, size=len(pairs))
    H = Hamiltonian()
    for i in range(f):
        H.add_term(1.0, [(i, FACTOR_POLY, (1.0, 0.0, 0.5, quartic))])
    for (i, j), cij in zip(pairs, c):
        H.add_term(cij, [(i, FACTOR_X), (j, FACTOR_X)])
    return H


# ----------------------------------------------------------------------------
# The five workloads (sizes as stated for each)
# ----------------------------------------------------------------------------

def cmf8_model(seed: int = 2024):
    """8 coupled HO modes, 8 HO-DVR points, 4 SPFs, displaced Gaussian
    (q0 = 1, p0 = 0) initial Hartree product; CMF dt = 0.1."""
    text = balanced_tree_text(8, 8, [4, 4, 4], omega=1.0, r0=0.0, wfr0=1.0, wfomega=1.0)
    return text, coupled_ho(8, seed)


def sweep64_model(seed: int = 64):
    """64-mode balanced binary tree, 16 DVR points, SPFs 8,10,11,13,14,16 by
    layer (bottom to the top's children); 64 one-mode + 2016 pair terms,
    c_ij ~ N(0, 0.01^2)."""
    text = balanced_tree_text(64, 16, [8, 10, 11, 13, 14, 16])
    return text, coupled_ho(64, seed, c_sigma=0.01, one_mode_matrix=True)


def cmf64_model(seed: int = 64):
    """The 64-mode sweep tree and Hamiltonian with displaced Gaussian leaves
    (q0 = 1, p0 = 0) for a Hartree-product CMF start."""
    text = balanced_tree_text(64, 16, [8, 10, 11, 13, 14, 16], wfr0=1.0)
    return text, coupled_ho(64, seed, c_sigma=0.01, one_mode_matrix=True)


def grad32_model(seed: int = 32):
    """32-mode anharmonic model: 8 DVR, SPF 8 at the bottom and 24 above."""
    text = balanced_tree_text(32, 8, [8, 24, 24, 24, 24])
    return text, coupled_quartic(32, seed)


def cmf128_model(seed: int = 128):
    """128-mode binary tree, 8 DVR, SPFs 8,14,21,27,34,41,48 toward the root,
    coupled-quartic SOP (128 + 8128 terms)."""
    text = balanced_tree_text(128, 8, [8, 14, 21, 27, 34, 41, 48])
    return text, coupled_quartic(128, seed)


def eig8_model(nstates: int = 3, seed: int = 7):
    """Imaginary-time eigenstates workload: 8 coupled HO modes with
    frequencies w_i = 1 + 0.25 i (an HO-DVR of 8 points at each mode's own
    frequency), pair couplings c_ij ~ N(0, 0.02^2), 4 SPFs per node and
    ``nstates`` states at the top.  Returns the tree text, the Hamiltonian and
    the force-constant matrix K (K_ii = w_i^2, K_ij = c_ij)."""
    f = 8
    w = 1.0 + 0.25 * np.arange(f)
    rng = np.random.default_rng(seed)
    H = Hamiltonian()
    for i in range(f):
        H.add_term(1.0, [(i, FACTOR_POLY, (1.0, 0.0, 0.5 * w[i] ** 2, 0.0))])
    K = np.diag(w ** 2)
    for i, j in _pairs(f):
        c = rng.normal(0.0, 0.02)
        K[i, j] = K[j, i] = c
        H.add_term(c, [(i, FACTOR_X), (j, FACTOR_X)])
    text = balanced_tree_text(f, 8, [4, 4, 4], omega=w, wfomega=w, nstates=nstates)
    return text, H, K


def microbench_shapes(count: int = 4096, seed: int = 4096):
    """Order-3 tensors with legs drawn from {8,16,24,32,48,64} and a random
    mode per tensor."""
    rng = np.random.default_rng(seed)
    legs = np.array([8, 16, 24, 32, 48, 64], dtype=np.int64)
    shapes = legs[rng.integers(0, 6, size=(count, 3))]
    modes = rng.integers(0, 3, size=count).astype(np.int64)
    return shapes, modes
