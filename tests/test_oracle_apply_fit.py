"""Pins of the oracle's SOP application fitted into a tree
(applyOperatorIteration, ApplyFirstOrder.cpp:43-147; fidelity,
TreeApplyOperator.cpp:22-35) by what the mathematics fixes, computed densely:

- with complete SPF sets the fit is the normalised H|Psi> itself (up to a
  global phase) and its fidelity |<fit|H Psi>|^2 equals |H Psi|^2;
- on a truncated tree the reported fidelity is |<fit|H Psi>|^2 of the dense
  vectors, it never exceeds |H Psi|^2 and it does not fall from one
  iteration to the next;
- a single-product operator is applied exactly (applyReference).
The operator is the SOP's bare products: the reference sums sop(i) without
the coefficients, so the test operators carry coefficient 1."""
import numpy as np
import pytest

import dense_ref as D
import oracle_lib as O
from qutree_amd import workloads as W


def unit(H):
    """The same products with coefficient 1."""
    U = W.Hamiltonian()
    for l in range(H.nterms):
        U.coeff.append(1.0)
        for f in range(H.term_off[l], H.term_off[l + 1]):
            U.fmode.append(H.fmode[f])
            U.fkind.append(H.fkind[f])
            U.fpool.append(H.fpool[f])
        U.term_off.append(len(U.fmode))
    U.pool = list(H.pool)
    return U


def setup(text, H, seed=5):
    m = O.OracleModel(text, H)
    psi = m.fill_random(seed, delta_lowest=False)
    grids = {}
    for mo in range(len([n for n in m.nodes if n["bottom"]])):
        x, kin, tr = m.leaf(mo)
        xr, Ur, kr = D.ho_dvr(len(x))
        grids[mo] = (xr, kr)
    dt = D.DenseTree(m, psi)
    return m, psi, dt.psi.ravel(), dt.h_psi(H, grids).ravel()


def test_complete_spfs_fit_is_exact():
    text = W.balanced_tree_text(4, 3, [3, 9])
    H = unit(W.coupled_ho(4, 17, c_sigma=0.2))
    m, psi, v, hv = setup(text, H)
    r = m.apply_fit(psi, 11)
    fit = D.DenseTree(m, r["psi"]).psi.ravel()
    nh = np.vdot(hv, hv).real
    assert abs(np.vdot(fit, fit).real - 1.0) < 1e-12
    assert abs(abs(np.vdot(fit, hv)) ** 2 - nh) < 1e-10 * nh
    assert abs(r["fidelities"][-1] - nh) < 1e-10 * nh


def contraction_sop(f, d, nterms, seed):
    """nterms products of two random d x d leaf matrices of spectral norm 1/2
    on mode pairs: |H| <= nterms / 4, so |H Psi|^2 < 1 and the iteration runs
    until the fidelity stops rising (the reference stops at once when
    1 - f < 1e-6, i.e. for every |H Psi| >= 1)."""
    rng = np.random.default_rng(seed)
    H = W.Hamiltonian()
    for t in range(nterms):
        i, j = rng.choice(f, 2, replace=False)
        mats = []
        for _ in range(2):
            A = rng.normal(size=(d, d)) + 1j * rng.normal(size=(d, d))
            mats.append(0.5 * A / np.linalg.norm(A, 2))
        H.add_term(1.0, [(int(i), W.FACTOR_MATRIX, mats[0]), (int(j), W.FACTOR_MATRIX, mats[1])])
    return H


@pytest.mark.parametrize("kind", ["f4", "f8"])
def test_truncated_fit_fidelity_is_the_dense_overlap(kind):
    if kind == "f4":
        text, H = W.balanced_tree_text(4, 3, [2, 3]), contraction_sop(4, 3, 4, 1)
    else:
        text, H = W.balanced_tree_text(8, 2, [2, 3, 3]), contraction_sop(8, 2, 4, 2)
    m, psi, v, hv = setup(text, H)
    r = m.apply_fit(psi, 3)
    f = r["fidelities"]
    fit = D.DenseTree(m, r["psi"]).psi.ravel()
    assert abs(np.vdot(fit, fit).real - 1.0) < 1e-12
    assert abs(abs(np.vdot(fit, hv)) ** 2 - f[-1]) < 1e-10 * f[-1]
    assert f[-1] <= np.vdot(hv, hv).real * (1 + 1e-12)
    assert np.all(np.diff(f) > -1e-12)
    assert len(f) >= 2


def test_single_product_is_applied_exactly():
    text = W.balanced_tree_text(4, 3, [2, 3])
    H = W.Hamiltonian()
    H.add_term(1.0, [(1, W.FACTOR_X), (2, W.FACTOR_KIN)])
    m, psi, v, hv = setup(text, H)
    r = m.apply_fit(psi, 1)
    got = D.DenseTree(m, r["psi"]).psi.ravel()
    assert np.max(np.abs(got - hv)) < 1e-13 * np.max(np.abs(hv))
    assert list(r["fidelities"]) == [1.0]
