"""GPU energy gradient dE/dA* (qtg_engine_gradient): parity with the oracle's
EnergyGradient on small trees, central finite differences of the GPU energy
on the 32-mode anharmonic model (8 DVR, SPF 8 then 24: nodes of 24 x 24 x 24)
at 64 seeded node entries with h = 1e-6, and gradient-descent steps that
lower the energy."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from canon import canonicalize_all, canonicalize_path
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def _models():
    return {
        "f4": (W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1)),
        "f8quartic": (W.balanced_tree_text(8, 2, [2, 3, 3]), W.coupled_quartic(8, 12, c_sigma=0.1)),
        "cmf8": W.cmf8_model(),
        "f8big": (W.balanced_tree_text(8, 8, [8, 24, 24]), W.coupled_quartic(8, 5, c_sigma=0.05)),
    }


@pytest.mark.parametrize("compress", [True, False])
@pytest.mark.parametrize("name", ["f4", "f8quartic", "cmf8", "f8big"])
def test_gradient_matches_oracle(name, compress):
    text, H = _models()[name]
    print(f"case gradient {name} compress={compress}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(21, delta_lowest=False)
    ref, e_ref = om.gradient(psi)
    eng = Q.Engine(text, H, compress=compress)
    eng.set_psi(psi)
    got, e = eng.gradient()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  max rel err {err:.3e}", flush=True)
    assert abs(e - e_ref) < 1e-12 * abs(e_ref)
    assert err < 1e-11


def _gpu_energy(eng, nodes, psi, a):
    q = canonicalize_path(psi.copy(), nodes, a)
    eng.set_psi(q)
    eng.derivative()
    return eng.energy()


def test_gradient_finite_differences_grad32():
    text, H = W.grad32_model()
    print("case gradient finite differences grad32 (64 entries, h = 1e-6)", flush=True)
    eng = Q.Engine(text, H)
    nodes = [dict(n, address=i) for i, n in enumerate(eng.nodes)]
    psi = eng.fill_random(32, delta_lowest=False)
    eng.set_psi(psi)
    g, E = eng.gradient()
    rng = np.random.default_rng(64)
    entries = rng.choice(eng.psi_size, size=64, replace=False)
    h = 1e-6
    worst = 0.0
    scale = np.max(np.abs(g))
    for e in entries:
        a = next(i for i, n in enumerate(nodes) if n["off"] <= e < n["off"] + int(np.prod(n["dims"])))
        fd = 0.0j
        for unit, w in ((1.0, 0.5), (1.0j, 0.5j)):
            p1, p2 = psi.copy(), psi.copy()
            p1[e] += h * unit
            p2[e] -= h * unit
            fd += w * (_gpu_energy(eng, nodes, p1, a) - _gpu_energy(eng, nodes, p2, a)) / (2 * h)
        worst = max(worst, abs(fd - g[e]))
    print(f"  E {E:.12f}  max |fd - g| {worst:.3e}  max |g| {scale:.3e}", flush=True)
    assert worst < 1e-6 * scale


def test_gradient_descent_lowers_energy():
    text, H = W.grad32_model()
    print("case gradient descent grad32 (20 steps)", flush=True)
    eng = Q.Engine(text, H)
    nodes = [dict(n, address=i) for i, n in enumerate(eng.nodes)]
    psi = eng.fill_random(32, delta_lowest=False)
    eta = 0.05
    energies = []
    for it in range(20):
        eng.set_psi(psi)
        g, E = eng.gradient()
        energies.append(E)
        psi = canonicalize_all(psi - eta * g, nodes)
    eng.set_psi(psi)
    eng.derivative()
    energies.append(eng.energy())
    print(f"  energies {energies[0]:.8f} -> {energies[-1]:.8f}", flush=True)
    assert all(b < a for a, b in zip(energies, energies[1:]))
