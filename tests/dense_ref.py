"""Independent dense (full configuration space) numpy restatement used to pin
the oracle's tree algorithms on small trees.  Nothing here follows the
reference's sweep structure: every quantity is computed from the full wave
function psi(x_0, ..., x_{f-1}) and its reduced objects."""
from __future__ import annotations

import numpy as np
from numpy.polynomial.hermite import hermgauss

from qutree_amd import workloads as W


def ho_dvr(d: int, omega: float = 1.0, r0: float = 0.0):
    """HO-DVR grid, FBR->DVR transformation (first row >= 0) and the kinetic
    matrix in the DVR, from the position matrix of the HO basis."""
    X = np.zeros((d, d))
    for n in range(d - 1):
        X[n + 1, n] = X[n, n + 1] = np.sqrt((n + 1) / (2.0 * omega))
    x, U = np.linalg.eigh(X)
    U = U * np.sign(U[0])
    # T = p^2/2 = -(omega/4) (a^+ - a)^2 with exact HO matrix elements
    # <n|T|n> = omega (2n+1)/4, <n|T|n+2> = -omega sqrt((n+1)(n+2))/4;
    # checked by the exact spectrum test in test_oracle_tree.py.
    K = np.zeros((d, d))
    for j in range(d):
        K[j, j] = (2 * j + 1) * omega / 4.0
        if j + 2 < d:
            K[j, j + 2] = K[j + 2, j] = -np.sqrt((j + 1) * (j + 2)) * omega / 4.0
    kin = U.T @ K @ U
    return x + r0, U, kin


def leaf_matrix(kind, pool_vals, x, kin, p=None):
    d = len(x)
    if kind == W.FACTOR_X:
        return np.diag(x).astype(np.complex128)
    if kind == W.FACTOR_X2:
        return np.diag(x * x).astype(np.complex128)
    if kind == W.FACTOR_KIN:
        return kin.astype(np.complex128)
    if kind == W.FACTOR_ID:
        return np.eye(d, dtype=np.complex128)
    if kind == W.FACTOR_POLY:
        aT, a1, a2, a4 = pool_vals
        return (aT * kin + np.diag(a1 * x + a2 * x ** 2 + a4 * x ** 4)).astype(np.complex128)
    raise ValueError(kind)


def term_factors(H: W.Hamiltonian, l, grids):
    """[(mode, matrix)] of term l in application order."""
    out = []
    for f in range(H.term_off[l], H.term_off[l + 1]):
        mode, kind = H.fmode[f], H.fkind[f]
        x, kin = grids[mode]
        if kind == W.FACTOR_MATRIX:  # d x d column-major, interleaved complex, at the pool offset
            d = len(x)
            raw = np.asarray(H.pool[2 * H.fpool[f]: 2 * H.fpool[f] + 2 * d * d])
            out.append((mode, (raw[0::2] + 1j * raw[1::2]).reshape(d, d, order="F")))
            continue
        vals = H.pool[2 * H.fpool[f]: 2 * H.fpool[f] + 4] if kind == W.FACTOR_POLY else None
        out.append((mode, leaf_matrix(kind, vals, x, kin)))
    return out


class DenseTree:
    """Node tensors of an oracle model, expanded to dense basis functions."""

    def __init__(self, model, psi):
        self.m = model
        self.nodes = model.nodes
        self.A = [model.node_tensor(psi, a) for a in range(model.nnodes)]
        self.modes = {}
        self.phi = {}
        for n in self.nodes:  # address order = children first
            a = n["address"]
            if n["bottom"]:
                self.modes[a] = [n["mode"]]
                self.phi[a] = self.A[a].T.copy()  # (n, d)
            else:
                ms = []
                T = self.A[a]
                for k, c in enumerate(n["children"]):
                    ms += self.modes[c]
                # contract children one at a time, child 0 first
                cur = T
                for k, c in enumerate(n["children"]):
                    cur = np.tensordot(cur, self.phi[c], axes=([0], [0]))
                # cur axes: (n, child0 grid axes..., child1 grid axes...)
                self.phi[a] = cur
                self.modes[a] = ms
        top = len(self.nodes) - 1
        self.f = len(self.modes[top])
        assert self.modes[top] == list(range(self.f)), "dense helper needs modes in leaf order"
        self.psi = self.phi[top][0]

    def apply_factors(self, vec, factors, axes_of_mode):
        """apply [(mode, matrix)] to a dense array whose axis for mode m is axes_of_mode[m]."""
        for mode, M in factors:
            if mode not in axes_of_mode:
                continue
            ax = axes_of_mode[mode]
            vec = np.moveaxis(np.tensordot(M, vec, axes=([1], [ax])), 0, ax)
        return vec

    def h_psi(self, H, grids, terms=None):
        full_axes = {m: m for m in range(self.f)}
        out = np.zeros_like(self.psi)
        for l in range(H.nterms) if terms is None else terms:
            out += H.coeff[l] * self.apply_factors(self.psi, term_factors(H, l, grids), full_axes)
        return out

    def sub_axes(self, a):
        return {m: i + 1 for i, m in enumerate(self.modes[a])}  # in phi[a], axis 0 is the state

    def hmat(self, H, grids, l, a):
        Phi = self.phi[a]
        hPhi = self.apply_factors(Phi, term_factors(H, l, grids), self.sub_axes(a))
        n = Phi.shape[0]
        return Phi.reshape(n, -1).conj() @ hPhi.reshape(n, -1).T

    def single_hole(self, a):
        """G[j, comp axes] = <Phi_a^j | psi> over the subtree's modes."""
        ms = self.modes[a]
        Phi = self.phi[a]
        return np.tensordot(Phi.conj(), self.psi, axes=(list(range(1, len(ms) + 1)), ms))

    def comp_axes(self, a):
        comp = [m for m in range(self.f) if m not in self.modes[a]]
        return comp, {m: i + 1 for i, m in enumerate(comp)}

    def rho(self, a):
        G = self.single_hole(a)
        n = G.shape[0]
        G = G.reshape(n, -1)
        return G.conj() @ G.T

    def hole(self, H, grids, l, a):
        G = self.single_hole(a)
        comp, ax = self.comp_axes(a)
        hG = self.apply_factors(G, term_factors(H, l, grids), ax)
        n = G.shape[0]
        return G.reshape(n, -1).conj() @ hG.reshape(n, -1).T

    def apply_node(self, H, grids, a, active_terms):
        """<chi^alpha G_k | sum_l H_l | psi> over the terms active at node a."""
        hpsi = self.h_psi(H, grids, active_terms)
        n = self.nodes[a]
        ms = self.modes[a]
        comp = [m for m in range(self.f) if m not in ms]
        if n["parent"] < 0:
            Y = hpsi[None, ...]
        else:
            G = self.single_hole(a)
            # contract conj(G_k) over complement axes
            Y = np.tensordot(G.conj(), hpsi, axes=(list(range(1, len(comp) + 1)), comp))
        # Y: (k, sub axes...)
        if n["bottom"]:
            return Y.T.copy()  # (d, k)
        cur = Y
        for c in n["children"]:
            nc = len(self.modes[c])
            cur = np.tensordot(cur, self.phi[c].conj(), axes=(list(range(1, 1 + nc)), list(range(1, 1 + nc))))
        # cur: (k, alpha0, alpha1, ...)
        return np.moveaxis(cur, 0, -1)
