"""GPU parity on the full-size workloads:
- the node microbench at its full size (4096 order-3 tensors) against the
  oracle, every tensor;
- the 128-mode sweep and the first CMF steps of the 64-mode tree (the
  workload behind the CMF steps/s figure) against oracle fixtures written by
  tests/golden/make_oracle_fixtures.py (the oracle needs minutes for them);
- the precision run of the 128-mode tree: CMF with complex64 bulk
  contractions (split form) against the complex128 device run, eps = 1e-8."""
import os

import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b)))


def test_microbench_all_4096_tensors():
    """matrixTensor and contraction on all 4096 tensors of the microbench
    (legs from {8,16,24,32,48,64}), complex128, each tensor within 1e-12 of
    the oracle."""
    from test_gpu_node import _microbench_problem

    count = 4096
    print(f"case microbench full count={count}", flush=True)
    shapes, modes, offT, offM, A, B, M, mp, hc = _microbench_problem(count, 4096)
    Cref, Sref = np.zeros_like(A), np.zeros_like(M)
    cr, sr = O.cview(Cref), O.cview(Sref)
    O.lib().orc_microbench_batch(count, np.ascontiguousarray(shapes.ravel()), modes, offT, offM, O.cview(A),
                                 O.cview(B), O.cview(M), cr, sr)
    Cref, Sref = O.as_complex(cr), O.as_complex(sr)
    dA, dB, dM = (Q.DeviceBuffer.from_array(x) for x in (A, B, M))
    del B
    dC, dS = Q.DeviceBuffer(A.nbytes), Q.DeviceBuffer(M.nbytes)
    Q.Plan("mode_product", mp).execute(dC, dM, dA)
    Q.Plan("hole", hc).execute(dS, dA, dB)
    Q.synchronize()
    got_C, got_S = dC.download(np.complex128), dS.download(np.complex128)
    errC = np.maximum.reduceat(np.abs(got_C - Cref), offT) / np.maximum.reduceat(np.abs(Cref), offT)
    errS = np.maximum.reduceat(np.abs(got_S - Sref), offM) / np.maximum.reduceat(np.abs(Sref), offM)
    print(f"  worst tensor: matrixTensor {errC.max():.2e}, contraction {errS.max():.2e}", flush=True)
    assert errC.max() < 1e-12 and errS.max() < 1e-12


def tree_overlap(eng, bra, ket):
    """<bra|ket> of two tree states (engine layout): S matrices bottom-up
    (MatrixTreeFunctions_Implementation.h:16-56), numpy."""
    S = {}
    for a, n in enumerate(eng.nodes):  # addresses are post-order: children first
        K = eng.node_tensor(ket, a)
        for c_idx, c in enumerate(n["children"]):
            K = np.moveaxis(np.tensordot(S[c], K, axes=([1], [c_idx])), 0, c_idx)
        B = eng.node_tensor(bra, a)
        nl = B.shape[-1]
        S[a] = B.reshape(-1, nl, order="F").conj().T @ K.reshape(-1, nl, order="F")
    return S[len(eng.nodes) - 1]


def fixture(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def node_checks(eng, d, psi, fx, tol):
    top = eng.nnodes - 1
    norms = np.array([np.linalg.norm(eng.node_tensor(d, a)) for a in range(eng.nnodes)])
    ovl = np.array([np.vdot(eng.node_tensor(psi, a), eng.node_tensor(d, a)) for a in range(eng.nnodes)])
    assert rel(norms, fx["norms"]) < tol
    # <A_v|dPsi_v> vanishes below the top ((1 - P) projects it out): compare
    # on the scale of the largest node derivative
    assert np.max(np.abs(ovl - fx["overlaps"])) < tol * np.max(fx["norms"])
    assert rel(eng.node_tensor(d, top).ravel(order="F"), fx["top"]) < tol


def test_sweep128_matches_oracle_fixture():
    """One sweep of the 128-mode tree (cfg5 tree and SOP, 8256 terms),
    complex128: energy, top node, four bottom nodes and per-node checksums
    of the derivative against the oracle fixture, 1e-11."""
    fx = fixture("sweep128.npz")
    text, H = W.cmf128_model()
    print("case sweep128 complex128 vs fixture", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(int(fx["seed"]), delta_lowest=False)
    eng.set_psi(psi)
    eng.derivative()
    d = eng.get(0)
    e = eng.energy()
    print(f"  energy {e:.15f} oracle {float(fx['energy']):.15f}", flush=True)
    assert abs(e - float(fx["energy"])) < 1e-12 * abs(float(fx["energy"]))
    node_checks(eng, d, psi, fx, 1e-11)
    # the bottom nodes hold all 8 SPFs of their 8-point leaf, so (1 - P)
    # leaves only rounding there: compare on the scale of the largest node
    scale = float(np.max(fx["norms"]))
    for a in range(4):
        assert np.max(np.abs(eng.node_tensor(d, a).ravel(order="F") - fx[f"node{a}"])) < 1e-11 * scale


@pytest.mark.parametrize("precision,tol", [(2, 1e-8), (1, 1e-6)])
def test_sweep128_reduced_precision_energy(precision, tol):
    """The complex64 forms on the 128-mode sweep against the FP64 oracle
    fixture: the split form within 1e-8 relative energy error, the all-FP32
    bulk within 1e-6."""
    fx = fixture("sweep128.npz")
    text, H = W.cmf128_model()
    print(f"case sweep128 precision={precision} vs fixture", flush=True)
    eng = Q.Engine(text, H, precision=precision)
    psi = eng.fill_random(int(fx["seed"]), delta_lowest=False)
    eng.set_psi(psi)
    eng.derivative()
    e = eng.energy()
    err = abs(e - float(fx["energy"])) / abs(float(fx["energy"]))
    print(f"  energy rel err {err:.2e}", flush=True)
    assert err < tol


def test_cmf64_first_steps_match_oracle_fixture():
    """The CMF workload of the steps/s figure (64-mode tree, displaced
    Hartree start, dt = 0.1, cmf 1e-4, bs 1e-5), its first CMF steps:
    identical step decisions, energies and autocorrelations at every output
    time, final top node and per-node checksums against the oracle.  A
    Hartree start sits on the ill-conditioned floor measured in
    tests/test_oracle_floor.py, so the bar is 1e-9 (energies) and 1e-6 (top
    node), not 1e-12."""
    fx = fixture("cmf64_cmf.npz")
    text, H = W.cmf64_model()
    t0, tend, dt, out, acc_cmf, acc_bs = (float(x) for x in fx["args"])
    print(f"case cmf64 first {int(fx['steps'])} steps vs fixture", flush=True)
    eng = Q.Engine(text, H)
    psi0 = eng.fill_random(int(fx["seed"]), delta_lowest=True)
    r = Q.CMF(eng).integrate(psi0, t0, tend, dt, out, acc_cmf=acc_cmf, acc_bs=acc_bs, max_steps=int(fx["steps"]))
    stats = [r["steps"], r["rejected"], r["builds"], r["derivs"]]
    print(f"  stats {stats} oracle {list(fx['stats'])}", flush=True)
    assert stats == [int(x) for x in fx["stats"]]
    assert len(r["times"]) == len(fx["times"])
    assert np.allclose(r["times"], fx["times"], rtol=0, atol=1e-12)
    de = np.max(np.abs(r["energies"] - fx["energies"]) / np.abs(fx["energies"]))
    da = np.max(np.abs(r["autocorr"] - fx["autocorr"]))
    top = eng.nnodes - 1
    dtop = np.max(np.abs(eng.node_tensor(r["psi"], top).ravel(order="F") - fx["top"]))
    norms = np.array([np.linalg.norm(eng.node_tensor(r["psi"], a)) for a in range(eng.nnodes)])
    print(f"  max rel dE {de:.2e} max |da| {da:.2e} top {dtop:.2e} norms {rel(norms, fx['norms']):.2e}", flush=True)
    assert de < 1e-9 and da < 1e-9
    assert dtop < 1e-6
    assert rel(norms, fx["norms"]) < 1e-9
    # the final state's energy and autocorrelation
    eng.set_psi(r["psi"])
    eng.derivative()
    e_end = eng.energy()
    a_end = tree_overlap(eng, psi0, r["psi"])[0, 0]
    print(f"  final energy {e_end:.15f} oracle {float(fx['final_energy']):.15f}", flush=True)
    assert abs(e_end - float(fx["final_energy"])) < 1e-9 * abs(float(fx["final_energy"]))
    assert abs(a_end - complex(fx["final_autocorr"])) < 1e-9


def test_cmf128_split_precision_against_complex128():
    """cfg5's precision question on the device: CMF on the 128-mode tree
    (8256 coupled-quartic terms) at dt = 0.05 with rho regularised at
    eps = 1e-8, complex64 bulk contractions with FP64 for the large terms,
    the densities, rho^-1, the projector and step control, against the
    complex128 run from the same random state: energies and
    autocorrelations within 1e-8 at every output time."""
    text, H = W.cmf128_model()
    print("case cmf128 split vs complex128, eps 1e-8", flush=True)
    res = {}
    for prec in (0, 2):
        eng = Q.Engine(text, H, precision=prec, rho_eps=1e-8)
        psi = eng.fill_random(128, delta_lowest=False)
        res[prec] = Q.CMF(eng).integrate(psi, 0.0, 0.25, 0.05, 0.05)
        print(f"  precision {prec}: steps {res[prec]['steps']} rejected {res[prec]['rejected']} outputs "
              f"{len(res[prec]['times'])}", flush=True)
        del eng
    a, b = res[2], res[0]
    assert len(a["times"]) == len(b["times"]) == 6
    de = np.max(np.abs(a["energies"] - b["energies"]) / np.abs(b["energies"]))
    da = np.max(np.abs(a["autocorr"] - b["autocorr"]))
    print(f"  max rel dE {de:.2e} max |da| {da:.2e}", flush=True)
    assert de < 1e-8 and da < 1e-8
