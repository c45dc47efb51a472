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


@pytest.mark.parametrize("start", ["hartree", "random"])
def test_cmf8_matches_oracle(start):
    text, H = W.cmf8_model()
    print(f"case cmf cmf8 {start}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(2024, delta_lowest=(start == "hartree"))
    om, eng, ref, got = run_both(text, H, psi, 1.0, 0.1, 0.1)
    de, da = report(ref, got)
    assert len(got["times"]) == len(ref["times"]) == 11
    assert np.allclose(got["times"], ref["times"], rtol=0, atol=1e-12)
    assert got["steps"] == ref["steps"] and got["rejected"] == ref["rejected"]
    assert got["derivs"] == ref["derivs"]
    # A Hartree start leaves the unoccupied SPFs with rho eigenvalues ~0, so
    # rho^-1 = 1/(lambda + 1e-10 e^{-lambda/1e-10}) multiplies their rounding
    # noise by ~1e10 in both implementations; energies and autocorrelation
    # then agree to ~1e-10 instead of ~1e-14 (random, full-rank starts), and
    # how close depends on the summation order of each contraction.
    assert de < (1e-9 if start == "hartree" else 1e-12)
    assert da < (1e-9 if start == "hartree" else 1e-12)


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
