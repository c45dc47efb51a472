"""GPU parity of imaginary-time eigenstates (Eigenstates.cpp:8-103 through
qtg_cmf_eigenstates) and of the engine on trees whose top node carries several
states, against the FP64 oracle on the same seeded trees, Hamiltonians and
states: levels, propagator energies, convergence overlaps, step statistics
and the final node tensors (each top-node state up to its phase)."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def node_rel(eng, a_vec, b_vec):
    floor = 1e-6 * np.max(np.abs(b_vec))
    worst = 0.0
    for a in range(eng.nnodes):
        x, y = eng.node_tensor(a_vec, a), eng.node_tensor(b_vec, a)
        worst = max(worst, np.max(np.abs(x - y)) / max(floor, np.max(np.abs(y))))
    return worst


@pytest.mark.parametrize("path", [1, 0])
def test_derivative_multi_state_matches_oracle(path):
    """Derivative and Apply of every node with a three-state top node: the
    top node's columns evolve under H and the densities sum over the states."""
    text, H, _ = W.eig8_model(3)
    print(f"case eigen derivative eig8 nstates=3 path={path}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(3, delta_lowest=False)
    ref = om.derivative(psi)
    app_ref = om.apply_all(psi)
    eng = Q.Engine(text, H, path=path)
    eng.set_psi(psi)
    eng.derivative()
    da, dd = node_rel(eng, eng.get(1), app_ref), node_rel(eng, eng.get(0), ref)
    print(f"  apply {da:.2e} derivative {dd:.2e}", flush=True)
    assert da < 1e-11
    assert dd < 1e-10


def compare_states(eng, got, ref):
    top = eng.nnodes - 1
    worst = 0.0
    for a in range(eng.nnodes):
        x, y = eng.node_tensor(got, a), eng.node_tensor(ref, a)
        if a == top:
            x = x.reshape(-1, x.shape[-1], order="F").copy()
            y = y.reshape(-1, y.shape[-1], order="F")
            for j in range(x.shape[1]):
                ov = np.vdot(x[:, j], y[:, j])
                x[:, j] *= ov / abs(ov)
        worst = max(worst, np.max(np.abs(x - y)) / np.max(np.abs(y)))
    print(f"  final state max rel {worst:.2e}", flush=True)
    return worst


@pytest.mark.parametrize("ns,iters,out", [(3, 3, 2.0), (1, 2, 4.0), (3, 0, 2.0)])
def test_eigenstates_match_oracle(ns, iters, out):
    text, H, K = W.eig8_model(ns)
    print(f"case eigenstates eig8 nstates={ns} iterations={iters} out={out}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(11, delta_lowest=False)
    kw = dict(acc_cmf=1e-7, acc_bs=1e-8)
    ref = om.eigenstates(psi, 0.05, out, iters, **kw)
    eng = Q.Engine(text, H)
    got = Q.CMF(eng).eigenstates(psi, 0.05, out, iters, **kw)
    de = np.max(np.abs(got["energies"] - ref["energies"]) / np.abs(ref["energies"]))
    print(f"  steps {got['steps']}/{ref['steps']} rejected {got['rejected']}/{ref['rejected']} builds "
          f"{got['builds']}/{ref['builds']} derivs {got['derivs']}/{ref['derivs']}  max rel dE {de:.2e}", flush=True)
    if iters:
        dp = np.max(np.abs(got["propagator"] - ref["propagator"]) / np.abs(ref["propagator"]))
        dc = np.max(np.abs(got["converged"] - ref["converged"]))
        print(f"  propagator {dp:.2e} converged {dc:.2e} dt {got['dt_end']:.6g}/{ref['dt_end']:.6g}", flush=True)
    ds = compare_states(eng, got["psi"], ref["psi"])
    # CMF decisions (steps, rejections, builds) are identical.  As the state
    # relaxes, the SPFs it no longer occupies get natural populations near
    # zero, and rho^-1 (regularised at 1e-10) multiplies the rounding of both
    # implementations on them by up to 1e10.  Over hundreds of steps a few
    # node Bulirsch-Stoer error tests then go the other way: the count of node
    # derivative evaluations changes by ~1e-3 relative, the CMF error
    # estimates (and so the step size carried out) by up to ~1e-4, the nodes
    # by a few 1e-8 (acc_cmf = 1e-7, acc_bs = 1e-8); the levels stay within
    # 1e-12.
    assert (got["steps"], got["rejected"], got["builds"]) == (ref["steps"], ref["rejected"], ref["builds"])
    assert abs(got["derivs"] - ref["derivs"]) <= 1e-2 * ref["derivs"]
    assert de < 1e-10
    if iters:
        assert dp < 1e-9
        assert dc < 1e-9
        assert abs(got["dt_end"] - ref["dt_end"]) < 1e-3 * ref["dt_end"]
    assert ds < 1e-7


def test_cmf_multi_state_matches_oracle():
    """Real-time CMF with three states at the top: energies h(0,0)/S(0,0) and
    autocorrelations of the first state, and the step statistics."""
    text, H, _ = W.eig8_model(3)
    print("case cmf eig8 nstates=3 real time", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(4, delta_lowest=False)
    ref = om.cmf(psi, 0.0, 0.3, 0.1, 0.1)
    eng = Q.Engine(text, H)
    got = Q.CMF(eng).integrate(psi, 0.0, 0.3, 0.1, 0.1)
    de = np.max(np.abs(got["energies"] - ref["energies"]) / np.abs(ref["energies"]))
    da = np.max(np.abs(got["autocorr"] - ref["autocorr"]))
    print(f"  outputs {len(got['times'])} steps {got['steps']}/{ref['steps']} derivs {got['derivs']}/"
          f"{ref['derivs']}  max rel dE {de:.2e}  max |da| {da:.2e}", flush=True)
    assert len(got["times"]) == len(ref["times"]) == 4
    assert got["steps"] == ref["steps"] and got["derivs"] == ref["derivs"]
    assert de < 1e-12
    assert da < 1e-12


def test_gradient_refuses_several_states():
    text, H, _ = W.eig8_model(3)
    eng = Q.Engine(text, H)
    eng.set_psi(eng.fill_random(1, delta_lowest=False))
    with pytest.raises(Q.QtgError, match="one state"):
        eng.gradient()
