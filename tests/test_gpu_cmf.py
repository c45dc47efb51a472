"""GPU parity of the CMF integrator (CMFIntegrator::Integrate with per-node
Bulirsch-Stoer, CMF error control and orthogonal()) against the FP64 oracle:
same seeded tree, Hamiltonian and initial state; energies and
autocorrelations at every output time, step statistics and final node
tensors."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def run_both(text, H, psi, t_end, dt, out, **kw):
    om = O.OracleModel(text, H)
    ref = om.cmf(psi, 0.0, t_end, dt, out, **kw)
    eng = Q.Engine(text, H)
    got = Q.CMF(eng).integrate(psi, 0.0, t_end, dt, out, **kw)
    return om, eng, ref, got


def report(ref, got):
    de = np.max(np.abs(got["energies"] - ref["energies"]) / np.abs(ref["energies"]))
    da = np.max(np.abs(got["autocorr"] - ref["autocorr"]))
    print(f"  outputs {len(got['times'])} steps {got['steps']}/{ref['steps']} rejected {got['rejected']}/"
          f"{ref['rejected']} derivs {got['derivs']}/{ref['derivs']} ticks {got['ticks']}  "
          f"max rel dE {de:.2e}  max |da| {da:.2e}", flush=True)
    return de, da


def rho_weighted_diff(om, x, y):
    """Largest LayerInterface::Error-style node difference
    sqrt(sum conj(D) rho D / tr rho) over the non-top nodes
    (src/TreeClasses/LayerInterface.cpp:6-31), rho of state x."""
    om.build(x)
    worst = 0.0
    for k in range(om.nnodes - 1):
        A, B = om.node_tensor(x, k), om.node_tensor(y, k)
        n = A.shape[-1]
        D = (A - B).reshape(-1, n, order="F")
        rho = om.rho(k)
        e = np.real(np.trace(D.conj().T @ D @ rho.T)) / np.real(np.trace(rho))
        worst = max(worst, float(np.sqrt(abs(e))))
    return worst


def diffs(om, a, b):
    top = om.nnodes - 1
    return dict(de=float(np.max(np.abs(a["energies"] - b["energies"]) / np.abs(b["energies"]))),
                da=float(np.max(np.abs(a["autocorr"] - b["autocorr"]))),
                node=rho_weighted_diff(om, b["psi"], a["psi"]),
                top=float(np.max(np.abs(om.node_tensor(a["psi"], top) - om.node_tensor(b["psi"], top)))))


@pytest.mark.parametrize("start", ["hartree", "random"])
def test_cmf8_matches_oracle(start):
    """cfg1 at its full horizon: 8 coupled HO modes, dt = 0.1, outputs every
    0.1 up to t = 10 (101 outputs).  Energies, autocorrelations, the
    rho-weighted node differences and the top node are compared with the
    oracle.  The tolerance is set by the oracle itself: the same oracle
    source built at -O2 without fast-math (another summation order) drifts
    from the -O3 -ffast-math build by a floor that a third correct FP64
    implementation cannot beat.  From a random (full-rank) start that floor
    is ~1e-14; from the Hartree start, whose unoccupied SPFs have density
    eigenvalues ~0 that rho^-1 = 1/(l + 1e-10 e^{-l/1e-10}) amplifies by
    ~1e10, it is ~2e-10 in the energy and ~3e-10 in the autocorrelation.
    The GPU must stay within 3 times the measured floor (and never needs to
    beat 1e-12)."""
    text, H = W.cmf8_model()
    print(f"case cmf cmf8 {start} t=10", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(2024, delta_lowest=(start == "hartree"))
    ref = om.cmf(psi, 0.0, 10.0, 0.1, 0.1)
    with O.strict_build():
        om2 = O.OracleModel(text, H)
        ref2 = om2.cmf(psi, 0.0, 10.0, 0.1, 0.1)
        del om2
    eng = Q.Engine(text, H)
    got = Q.CMF(eng).integrate(psi, 0.0, 10.0, 0.1, 0.1)
    report(ref, got)
    assert len(got["times"]) == len(ref["times"]) == len(ref2["times"]) == 101
    assert np.allclose(got["times"], ref["times"], rtol=0, atol=1e-12)
    assert got["steps"] == ref["steps"] and got["rejected"] == ref["rejected"]
    assert got["derivs"] == ref["derivs"]
    floor = diffs(om, ref2, ref)
    gpu = diffs(om, got, ref)
    print(f"  floor (oracle -O2 vs -O3 -ffast-math) {floor}", flush=True)
    print(f"  gpu vs oracle                          {gpu}", flush=True)
    for k in ("de", "da", "node", "top"):
        assert gpu[k] <= max(3.0 * floor[k], 1e-12), k
    if start == "random":
        assert floor["de"] < 1e-13 and gpu["de"] < 1e-12 and gpu["node"] < 1e-11


def test_cmf_f16_random_short():
    text = W.balanced_tree_text(16, 8, [4, 6, 6, 8])
    H = W.coupled_ho(16, 7, c_sigma=0.05)
    print("case cmf f16 random", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(16, delta_lowest=False)
    om, eng, ref, got = run_both(text, H, psi, 0.3, 0.1, 0.1)
    de, da = report(ref, got)
    assert got["steps"] == ref["steps"]
    assert de < 1e-10
    assert da < 1e-10
    # final tensors: top node and the rho-weighted node differences
    top = om.nnodes - 1
    x, y = om.node_tensor(got["psi"], top), om.node_tensor(ref["psi"], top)
    assert np.max(np.abs(x - y)) < 1e-9 * np.max(np.abs(y))


def test_cmf_energy_conservation_and_norm():
    # CMF with tight tolerances conserves <H> to ~1e-7 over t = 1 (at the
    # default 1e-4 / 1e-5 the method itself drifts by ~1e-4 per unit time, the
    # oracle identically); the norm of the top node stays 1.
    text, H = W.cmf8_model()
    print("case cmf cmf8 conservation t=1 tight", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(2024, delta_lowest=True)
    got = Q.CMF(eng).integrate(psi, 0.0, 1.0, 0.1, 0.25, acc_cmf=1e-7, acc_bs=1e-8)
    e = got["energies"]
    print(f"  energies {e[0]:.12f} .. {e[-1]:.12f}  steps {got['steps']}", flush=True)
    assert np.max(np.abs(e - e[0])) < 1e-6 * abs(e[0])
    assert abs(got["autocorr"][0] - 1.0) < 1e-12
    top = eng.nnodes - 1
    a = eng.node_tensor(got["psi"], top)
    assert abs(np.vdot(a, a).real - 1.0) < 1e-7  # BS tolerance 1e-8 per step, 267 steps


@pytest.mark.parametrize("precision,tol", [(2, 1e-8), (1, 1e-6)])
def test_cmf8_reduced_precision_against_oracle(precision, tol):
    """CMF with complex64 bulk contractions (FP64 integrator, projector, rho^-1):
    the split form (large-coefficient terms in FP64) stays within 1e-8 of the
    FP64 trajectory, the all-FP32 bulk within 1e-6."""
    text, H = W.cmf8_model()
    print(f"case cmf cmf8 precision={precision}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(2024, delta_lowest=False)
    ref = om.cmf(psi, 0.0, 1.0, 0.1, 0.1)
    eng = Q.Engine(text, H, precision=precision)
    got = Q.CMF(eng).integrate(psi, 0.0, 1.0, 0.1, 0.1)
    de, da = report(ref, got)
    assert got["steps"] == ref["steps"]
    assert de < tol and da < tol


def test_cmf8_global_memory_orthogonal_matches_oracle():
    """orthogonal() on node tensors too large for LDS runs in global memory
    with the square roots pushed by grouped mode products; forced here on
    the 8-mode tree, it must give the oracle's trajectory like the LDS
    kernel does (random start, t = 1)."""
    import os
    text, H = W.cmf8_model()
    print("case cmf cmf8 global-memory orthogonal", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(2024, delta_lowest=False)
    ref = om.cmf(psi, 0.0, 1.0, 0.1, 0.1)
    eng = Q.Engine(text, H)
    os.environ["QTG_ORTHO_GLOBAL"] = "1"
    try:
        cmf = Q.CMF(eng)
    finally:
        del os.environ["QTG_ORTHO_GLOBAL"]
    got = cmf.integrate(psi, 0.0, 1.0, 0.1, 0.1)
    de, da = report(ref, got)
    assert got["steps"] == ref["steps"] and got["derivs"] == ref["derivs"]
    assert de < 1e-12 and da < 1e-12
    top = om.nnodes - 1
    assert np.max(np.abs(om.node_tensor(got["psi"], top) - om.node_tensor(ref["psi"], top))) < 1e-11


@pytest.mark.parametrize("opts", [{"device_bs": 0}, {"device_bs": 0, "lookahead": 0}, {"device_bs": 0, "graph": 1},
                                  {"device_bs": 0, "fuse": 0}, {"device_bs": 0, "graph": 1, "fuse": 0}])
@pytest.mark.parametrize("model", ["cmf8", "cmf64"])
def test_cmf_tick_forms_equal_default(model, opts):
    """The default ticks run every node's Bulirsch-Stoer state machine on the
    device (one launch per tick boundary: post ops, accept test, next pre
    op).  The host state machine in each of its forms (ticks staged ahead
    with both outcomes of each accept test, or staged after the errors have
    reached the host; eager ticks or ticks replayed from HIP graphs; fused
    op sequences or one pass per op) runs the same arithmetic per element:
    energies, autocorrelations, statistics and the final state agree bit
    for bit, also across a second integrate call that continues from the
    first one's state."""
    text, H = getattr(W, model + "_model")()
    print(f"case cmf tick forms {model} {opts}", flush=True)
    runs = []
    for o in ({}, opts):
        eng = Q.Engine(text, H)
        psi = eng.fill_random(64 if model == "cmf64" else 8, delta_lowest=True)
        cmf = Q.CMF(eng)
        for k, v in o.items():
            cmf.set_option(k, v)
        a = cmf.integrate(psi, 0.0, 0.3, 0.1, 0.1)
        b = cmf.integrate(a["psi"], 0.3, 0.5, 0.1, 0.1)
        runs.append((a, b))
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x["energies"], y["energies"])
        assert np.array_equal(x["autocorr"], y["autocorr"])
        assert np.array_equal(x["psi"], y["psi"])
        for k in ("steps", "rejected", "builds", "derivs", "ticks"):
            assert x[k] == y[k], k


@pytest.mark.parametrize("device_bs", [1, 0])
def test_cmf_step_size_underflow_is_an_error(device_bs):
    """A Bulirsch-Stoer tolerance that no node error meets (0) makes every
    node reject until its step size underflows (dx + h == dx,
    BS_integrator.h:153-164, where the reference prints the message and
    exits the process).  The device integrators and the host state machine
    both end the call with that error; the condition is set from the host
    and every kernel exits as usual (about 14 rejection cycles of 312
    ticks)."""
    text, H = W.cmf8_model()
    print(f"case cmf step size underflow device_bs={device_bs}", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(8, delta_lowest=True)
    cmf = Q.CMF(eng)
    cmf.set_option("device_bs", device_bs)
    with pytest.raises(Q.QtgError, match="step size underflow"):
        cmf.integrate(psi, 0.0, 0.1, 0.1, 0.1, acc_cmf=1e-4, acc_bs=0.0)
