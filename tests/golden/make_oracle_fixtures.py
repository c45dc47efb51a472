"""Writes the oracle fixtures of the two large workloads (run once in the
build container; the FP64 oracle takes minutes on them, too long for a test):

cmf64_cmf.npz   CMF on the 64-mode tree (qutree_amd.workloads.cmf64_model:
                16 HO-DVR, SPF 8,10,11,13,14,16, displaced Hartree start from
                fill_random(64, delta_lowest=True)), dt = 0.1, outputs every
                0.1, cmf 1e-4, bs 1e-5, the first CMF_STEPS steps: output
                times, energies, autocorrelations, step / rejection / build /
                derivative counts, and the final state's energy,
                autocorrelation <Psi(0)|Psi>, top node and per-node checksums
                (Frobenius norm of each node tensor, and its overlap with the
                initial node tensor).
sweep128.npz    One H|Psi> sweep (Derivative, phase 1, eps 1e-10) on the
                128-mode tree (cmf128_model: 8 DVR, SPF 8,14,21,27,34,41,48,
                128 + 8128 coupled-quartic terms) at fill_random(128,
                delta_lowest=False): energy, the top node's derivative, the
                derivative of the first four nodes, and per-node checksums of
                the derivative (norm, <A_v|dPsi_v>).

The states are regenerated from their seeds by the engine's own
fill_random (the same mt19937 stream), so only checksums are stored.
Run from the repository root:  python tests/golden/make_oracle_fixtures.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))  # tests/
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))  # repository root

import oracle_lib as O  # noqa: E402
from qutree_amd import workloads as W  # noqa: E402

CMF_STEPS = 6
CMF_ARGS = dict(t0=0.0, tend=10.0, dt=0.1, out=0.1, acc_cmf=1e-4, acc_bs=1e-5)
CMF_SEED = 64
SWEEP_SEED = 128


def node_checksums(om, psi, ref):
    norms, ovl = [], []
    for a in range(om.nnodes):
        x, r = om.node_tensor(psi, a), om.node_tensor(ref, a)
        norms.append(np.linalg.norm(x))
        ovl.append(np.vdot(r, x))
    return np.array(norms), np.array(ovl)


def cmf64():
    text, H = W.cmf64_model()
    om = O.OracleModel(text, H)
    psi0 = om.fill_random(CMF_SEED, delta_lowest=True)
    t = time.time()
    r = om.cmf(psi0, CMF_ARGS["t0"], CMF_ARGS["tend"], CMF_ARGS["dt"], CMF_ARGS["out"], acc_cmf=CMF_ARGS["acc_cmf"],
               acc_bs=CMF_ARGS["acc_bs"], max_steps=CMF_STEPS)
    print(f"cmf64: {r['steps']} steps, {len(r['times'])} outputs, {time.time() - t:.1f} s", flush=True)
    norms, ovl = node_checksums(om, r["psi"], psi0)
    top = om.nnodes - 1
    om.build(r["psi"])  # Expectation reads the representation of the state it is given
    return dict(times=r["times"], energies=r["energies"], autocorr=r["autocorr"],
                stats=np.array([r["steps"], r["rejected"], r["builds"], r["derivs"]], np.int64),
                top=om.node_tensor(r["psi"], top).ravel(order="F"), norms=norms, overlaps=ovl,
                final_energy=np.array(om.energy(r["psi"])), final_autocorr=np.array(om.overlap(psi0, r["psi"])),
                seed=np.array(CMF_SEED), steps=np.array(CMF_STEPS),
                args=np.array([CMF_ARGS[k] for k in ("t0", "tend", "dt", "out", "acc_cmf", "acc_bs")]))


def sweep128():
    text, H = W.cmf128_model()
    om = O.OracleModel(text, H)
    psi = om.fill_random(SWEEP_SEED, delta_lowest=False)
    t = time.time()
    d = om.derivative(psi)
    e = om.energy(psi)
    print(f"sweep128: {time.time() - t:.1f} s", flush=True)
    norms, ovl = [], []
    for a in range(om.nnodes):
        x, A = om.node_tensor(d, a), om.node_tensor(psi, a)
        norms.append(np.linalg.norm(x))
        ovl.append(np.vdot(A, x))
    top = om.nnodes - 1
    out = dict(energy=np.array(e), top=om.node_tensor(d, top).ravel(order="F"), norms=np.array(norms),
               overlaps=np.array(ovl), seed=np.array(SWEEP_SEED))
    for a in range(4):
        out[f"node{a}"] = om.node_tensor(d, a).ravel(order="F")
    return out


def main():
    which = sys.argv[1:] or ["cmf64", "sweep128"]
    if "cmf64" in which:
        np.savez_compressed(os.path.join(HERE, "cmf64_cmf.npz"), **cmf64())
    if "sweep128" in which:
        np.savez_compressed(os.path.join(HERE, "sweep128.npz"), **sweep128())


if __name__ == "__main__":
    main()
