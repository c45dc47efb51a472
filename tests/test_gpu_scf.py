"""GPU parity of the Krylov-sweep SCF (qtg_scf_*, SCF.cpp:270-348) against
the FP64 oracle: the lowest Ritz value of every microstep and the final node
tensors, for the diagonalising and the imaginary-time updates, graph replay
against eager launches, and the reference's own SCF example input
(examples/mctdh/coupledHO.4.yaml: nIter 5, nKrylov 10)."""
import os

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import mctdh as M
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _model(kind):
    if kind == "f4":
        return W.balanced_tree_text(4, 3, [2, 3]), W.coupled_ho(4, 11, c_sigma=0.1)
    if kind == "f8":
        return W.balanced_tree_text(8, 2, [2, 3, 3]), W.coupled_quartic(8, 12, c_sigma=0.1)
    if kind == "cmf8":
        return W.cmf8_model()[:2]
    if kind == "quartic16":
        return W.balanced_tree_text(16, 6, [4, 6, 8, 8]), W.coupled_quartic(16, 5, c_sigma=0.01)
    if kind == "ho128":  # the reference's CoupledHarmonicOscillator.128.4.4 example tree
        return M.balanced_tree(128, 10, 4, 4, 0, 0.01822, 0.0, 0.0, 0.01822), M.coupled_ho(128)
    raise ValueError(kind)


def rho_weighted_diff(om, x, y):
    """Largest sqrt(sum conj(D) rho D / tr rho) over the non-top nodes, D the
    node difference and rho the density of state x: differences in SPF
    directions the state does not occupy (a free gauge of the QR moves when
    a node's density has near-zero eigenvalues) weigh nothing."""
    om.build(x)
    worst = 0.0
    for a in range(om.nnodes - 1):
        A, B = om.node_tensor(x, a), om.node_tensor(y, a)
        n = A.shape[-1]
        D = (A - B).reshape(-1, n, order="F")
        rho = om.rho(a)
        e = np.real(np.trace(D.conj().T @ D @ rho.T)) / np.real(np.trace(rho))
        worst = max(worst, float(np.sqrt(abs(e))))
    return worst


def measures(om, ref, got):
    """Energies of every microstep; the final state through gauge-invariant
    measures (fidelity of the whole wave function, density-weighted node
    differences) and the top node, which the last QR normalises."""
    de = float(np.max(np.abs(got["energies"] - ref["energies"])) / np.max(np.abs(ref["energies"])))
    ov = om.overlap(ref["psi"], got["psi"])
    nr, ng = om.overlap(ref["psi"], ref["psi"]).real, om.overlap(got["psi"], got["psi"]).real
    top = om.nnodes - 1
    return dict(de=de, infid=1.0 - abs(ov) ** 2 / (nr * ng), node=rho_weighted_diff(om, ref["psi"], got["psi"]),
                top=float(np.max(np.abs(om.node_tensor(got["psi"], top) - om.node_tensor(ref["psi"], top)))),
                raw=max(float(np.max(np.abs(om.node_tensor(got["psi"], a) - om.node_tensor(ref["psi"], a))))
                        for a in range(om.nnodes)))


def compare(om, ref, got, floor=None):
    """The GPU within 10 times the floor that the same oracle source built
    without fast-math (another summation order) shows against the default
    build, and never above 1e-10 in the energies, 1e-12 in the fidelity."""
    m = measures(om, ref, got)
    print(f"  microsteps {ref['energies'].size}  E {ref['energies'].ravel()[-1]:.14f}  " +
          "  ".join(f"{k} {v:.2e}" for k, v in m.items()), flush=True)
    if floor is not None:
        print("  floor   " + "  ".join(f"{k} {v:.2e}" for k, v in floor.items()), flush=True)
    f = floor or dict(de=0.0, node=0.0, top=0.0)
    assert m["de"] < 1e-10 and m["de"] <= max(10 * f["de"], 1e-12)
    assert m["infid"] < 1e-12
    assert m["node"] <= max(10 * f["node"], 1e-10)
    assert m["top"] <= max(10 * f["top"], 1e-10)


def oracle_pair(text, H, psi, *args, **kw):
    om = O.OracleModel(text, H)
    ref = om.scf(psi, *args, **kw)
    with O.strict_build():
        om2 = O.OracleModel(text, H)
        ref2 = om2.scf(psi, *args, **kw)
        del om2
    return om, ref, measures(om, ref, ref2)


@pytest.mark.parametrize("kind,k,iters", [("f4", 4, 3), ("f8", 3, 3), ("cmf8", 10, 2), ("quartic16", 6, 2)])
def test_scf_matches_oracle(kind, k, iters):
    text, H = _model(kind)
    print(f"case scf {kind} k={k} iters={iters}", flush=True)
    psi = O.OracleModel(text, H).fill_random(7, delta_lowest=False)
    om, ref, floor = oracle_pair(text, H, psi, iters, k)
    eng = Q.Engine(text, H)
    got = Q.SCF(eng, k).run(psi, iters)
    compare(om, ref, got, floor)


@pytest.mark.parametrize("beta", [0.0, 2.5])
def test_scf_imaginary_time_matches_oracle(beta):
    text, H = _model("f8")
    print(f"case scf f8 itp beta={beta}", flush=True)
    psi = O.OracleModel(text, H).fill_random(9, delta_lowest=True)
    om, ref, floor = oracle_pair(text, H, psi, 3, 4, n_itp=2, beta=beta)
    eng = Q.Engine(text, H)
    got = Q.SCF(eng, 4).run(psi, 3, n_itp=2, beta=beta)
    compare(om, ref, got, floor)


def test_scf_graph_replay_equals_eager():
    text, H = _model("cmf8")
    print("case scf cmf8 graph vs eager", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(3, delta_lowest=True)
    s1 = Q.SCF(eng, 6)
    a = s1.run(psi, 2)
    b = s1.run(psi, 2)  # the same graph replayed again
    s2 = Q.SCF(eng, 6)
    s2.set_option("graph", 0)
    c = s2.run(psi, 2)
    assert np.array_equal(a["energies"], b["energies"]) and np.array_equal(a["psi"], b["psi"])
    assert np.max(np.abs(a["energies"] - c["energies"])) == 0.0
    assert np.max(np.abs(a["psi"] - c["psi"])) == 0.0


def test_scf_reference_example_coupled_ho4():
    """examples/mctdh/coupledHO.4.yaml of the reference: 4 coupled HO modes
    (10 DVR points, 4 and 10 SPFs), nIter 5, nKrylov 10, started from
    TensorTree(rng, tree, delta_lowest = true).  The SCF energy falls to the
    ground state of the 4 normal modes (sum of omega_k / 2 over the normal
    modes of the ring, a closed form) within the basis truncation."""
    tree = ("1 -2\n  10 -2\n    4 -1\n      10 0 0\n    4 -1\n      10 0 1\n  10 -2\n    4 -1\n      10 0 2\n"
            "    4 -1\n      10 0 3\n" + "0.01822 0.0 0.0 0.0122\n" * 4)
    H = M.coupled_ho(4)
    print("case scf coupledHO.4 example", flush=True)
    psi = O.OracleModel(tree, H).fill_random(1234, delta_lowest=True)
    om, ref, floor = oracle_pair(tree, H, psi, 5, 10)
    eng = Q.Engine(tree, H)
    got = Q.SCF(eng, 10).run(psi, 5)
    compare(om, ref, got, floor)
    # V = 1/2 om^2 sum x_i^2 + lam^2 sum x_i x_{i+1}: Hessian om^2 on the
    # diagonal, lam^2 between ring neighbours (CoupledHO.cpp:7-66)
    lam, om_ = 2000.0 / M.CM, 4000.0 / M.CM
    K = om_ ** 2 * np.eye(4)
    for i in range(4):
        K[i, (i + 1) % 4] += lam ** 2
        K[(i + 1) % 4, i] += lam ** 2
    e0 = 0.5 * np.sum(np.sqrt(np.linalg.eigvalsh(K)))
    e = got["energies"].ravel()[-1]
    print(f"  final {e:.12f} closed form {e0:.12f} rel gap {(e - e0) / e0:.2e}", flush=True)
    assert e >= e0 * (1 - 1e-12) and e - e0 < 1e-6 * e0


def test_scf_rejects_bad_arguments():
    text, H = _model("f4")
    eng = Q.Engine(text, H)
    with pytest.raises(RuntimeError):
        Q.SCF(eng, 0)
    with pytest.raises(RuntimeError):
        Q.SCF(eng, 65)
    e32 = Q.Engine(text, H, precision=1)
    with pytest.raises(RuntimeError):
        Q.SCF(e32, 4)


def test_scf_ho128_reference_example():
    """The reference's 128-mode SCF example (CoupledHarmonicOscillator.128.4.4:
    SPF 4 + 4 per layer, Hartree start, nKrylov 10).  Its sweep is sensitive
    to rounding in the first iterations: the unoccupied SPF directions each QR
    completes arbitrarily become the next node's search space, and the two
    oracle builds (-O3 -ffast-math and -O2) already differ by ~4e-8 in the
    energies after one iteration (from a random start by ~4e-4).  Checked:
    after one iteration the GPU's energies lie within 30 times that floor of
    the oracle's; the energies never rise along the sweep; the last reported
    Ritz value is the energy of the returned state (the oracle's own energy
    function on it); after six iterations GPU and oracle agree to 1e-8 and
    lie above the closed-form ground state by less than 1e-6."""
    text, H = _model("ho128")
    print("case scf ho128 hartree", flush=True)
    psi = O.OracleModel(text, H).fill_random(128, delta_lowest=True)
    om, ref, floor = oracle_pair(text, H, psi, 1, 10)
    scf = Q.SCF(Q.Engine(text, H), 10)
    got = scf.run(psi, 1)
    m = measures(om, ref, got)
    print("  gpu   " + "  ".join(f"{k} {v:.2e}" for k, v in m.items()), flush=True)
    print("  floor " + "  ".join(f"{k} {v:.2e}" for k, v in floor.items()), flush=True)
    assert m["de"] <= 30 * floor["de"]
    e = got["energies"].ravel()
    assert np.all(np.diff(e[1:]) <= 1e-12 * abs(e[0]))
    om.build(got["psi"])
    e_state = om.energy(got["psi"])
    print(f"  last Ritz value {e[-1]:.15f}  energy of the state {e_state:.15f}", flush=True)
    assert abs(e_state - e[-1]) < 1e-10 * abs(e[-1])
    g6 = scf.run(psi, 6)["energies"][:, -1]
    r6 = om.scf(psi, 6, 10)["energies"][:, -1]
    lam, om_ = 2000.0 / M.CM, 4000.0 / M.CM
    e0 = 0.5 * float(np.sum(np.sqrt(om_ ** 2 + 2 * lam ** 2 * np.cos(2 * np.pi * np.arange(128) / 128))))
    print(f"  after 6 iterations gpu {g6[-1]:.13f} oracle {r6[-1]:.13f} closed form {e0:.13f}", flush=True)
    assert abs(g6[-1] - r6[-1]) < 1e-8 * r6[-1]
    assert e0 < g6[-1] < e0 * (1 + 1e-6)
