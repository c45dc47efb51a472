"""Pins the oracle's energy gradient dE/dA* (orthonormal gauge):

* central finite differences of the energy functional at every entry of
  small trees (the perturbed node is re-orthonormalised up to the top by QR,
  which leaves the state unchanged, so the oracle's energy applies), and
* rho^-1 (1 - P) of <Psi|Psi> dE/dA_v* equals minus the imaginary-time
  LayerDerivative (HamiltonianRepresentation.cpp:53-72 with phase -i), since
  the projector removes the mean-field and energy terms.
"""
import numpy as np
import pytest

import oracle_lib as O
from canon import canonicalize_path
from qutree_amd import workloads as W


def _models():
    return {
        "f4": (W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1)),
        "f8quartic": (W.balanced_tree_text(8, 2, [2, 3, 3]), W.coupled_quartic(8, 12, c_sigma=0.1)),
    }


def energy_of(om, psi, a):
    q = canonicalize_path(psi.copy(), om.nodes, a)
    om.build(q)
    return om.energy(q)


@pytest.mark.parametrize("name", ["f4", "f8quartic"])
def test_gradient_matches_finite_differences(name):
    text, H = _models()[name]
    om = O.OracleModel(text, H)
    psi = om.fill_random(7, delta_lowest=False)
    g, E = om.gradient(psi)
    om.build(psi)
    assert abs(E - om.energy(psi)) < 1e-12 * abs(E)
    h = 1e-5
    rng = np.random.default_rng(1)
    worst, scale = 0.0, np.max(np.abs(g))
    entries = rng.choice(om.psi_size, size=min(om.psi_size, 60), replace=False)
    for e in entries:
        a = next(n["address"] for n in om.nodes if n["off"] <= e < n["off"] + int(np.prod(n["dims"])))
        fd = 0.0j
        for unit, w in ((1.0, 0.5), (1.0j, 0.5j)):
            p1, p2 = psi.copy(), psi.copy()
            p1[e] += h * unit
            p2[e] -= h * unit
            # dE = 2 Re(conj(g) dA): d/dRe = 2 Re g, d/dIm = 2 Im g
            fd += w * (energy_of(om, p1, a) - energy_of(om, p2, a)) / (2 * h)
        worst = max(worst, abs(fd - g[e]))
    assert worst < 1e-7 * scale, (worst, scale)


@pytest.mark.parametrize("name", ["f4", "f8quartic"])
def test_projected_gradient_is_imaginary_time_derivative(name):
    text, H = _models()[name]
    om = O.OracleModel(text, H)
    psi = om.fill_random(8, delta_lowest=False)
    g, _ = om.gradient(psi)
    top = om.nodes[-1]
    Ntop = np.vdot(om.node_tensor(psi, top["address"]).ravel(), om.node_tensor(psi, top["address"]).ravel())
    d = om.derivative(psi, phase=-1j)
    for n in om.nodes:
        if n["parent"] < 0:
            continue
        a = n["address"]
        dims = n["dims"]
        N, k = int(np.prod(dims)), dims[-1]
        A = psi[n["off"]:n["off"] + N].reshape((N // k, k), order="F")
        X = (Ntop * g[n["off"]:n["off"] + N]).reshape((N // k, k), order="F")
        Y = X - A @ np.linalg.solve(A.conj().T @ A, A.conj().T @ X)
        got = Y @ om.rho_inv(a).T
        ref = -d[n["off"]:n["off"] + N].reshape((N // k, k), order="F")
        assert np.max(np.abs(got - ref)) < 1e-10 * max(1.0, np.max(np.abs(ref)))
