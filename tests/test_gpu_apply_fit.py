"""GPU parity of the fitted SOP application (qtg_apply_fit_*,
applyOperatorIteration of ApplyFirstOrder.cpp:43-147) against the FP64
oracle: the fidelity after every iteration and the fitted state (whole-state
fidelity between the two fits and the normalised top node), for contractive
operators that iterate to convergence, a full Hamiltonian (which the
reference's stopping rule ends after one iteration, its fidelity exceeding
1) and a single product, and graph replay against eager launches.  The
operator's coefficients are not applied, as in the reference."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
import test_oracle_apply_fit as TA
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def _model(kind):
    if kind == "contractive4":
        return W.balanced_tree_text(4, 3, [2, 3]), TA.contraction_sop(4, 3, 4, 1)
    if kind == "contractive8":
        return W.balanced_tree_text(8, 2, [2, 3, 3]), TA.contraction_sop(8, 2, 4, 2)
    if kind == "contractive16":
        return W.balanced_tree_text(16, 4, [3, 4, 6, 6]), TA.contraction_sop(16, 4, 12, 3)
    if kind == "cmf8":
        return W.cmf8_model()[:2]
    if kind == "single":
        H = W.Hamiltonian()
        H.add_term(1.0, [(1, W.FACTOR_X), (2, W.FACTOR_KIN)])
        return W.balanced_tree_text(4, 3, [2, 3]), H
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["contractive4", "contractive8", "contractive16", "cmf8", "single"])
def test_apply_fit_matches_oracle(kind):
    text, H = _model(kind)
    print(f"case apply_fit {kind}", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(21, delta_lowest=False)
    ref = om.apply_fit(psi, 5, 30)
    got = Q.ApplyFit(Q.Engine(text, H)).run(psi, 5, 30)
    f_ref, f_got = ref["fidelities"], got["fidelities"]
    ov = om.overlap(ref["psi"], got["psi"])
    nr, ng = om.overlap(ref["psi"], ref["psi"]).real, om.overlap(got["psi"], got["psi"]).real
    infid = 1.0 - abs(ov) ** 2 / (nr * ng)
    top = om.nnodes - 1
    dtop = float(np.max(np.abs(om.node_tensor(got["psi"], top) - om.node_tensor(ref["psi"], top))))
    print(f"  iterations {len(f_got)}/{len(f_ref)}  fidelity {f_ref[-1]:.14f}  "
          f"max rel df {np.max(np.abs(f_got - f_ref) / np.abs(f_ref)) if len(f_got) == len(f_ref) else -1:.2e}  "
          f"1-fidelity between the fits {infid:.2e}  top diff {dtop:.2e}", flush=True)
    assert len(f_got) == len(f_ref)
    assert np.max(np.abs(f_got - f_ref) / np.abs(f_ref)) < 1e-10
    assert infid < 1e-10
    if kind != "single":
        assert dtop < 1e-8


def test_apply_fit_graph_replay_equals_eager():
    text, H = _model("contractive8")
    print("case apply_fit graph vs eager", flush=True)
    eng = Q.Engine(text, H)
    psi = eng.fill_random(4, delta_lowest=False)
    a = Q.ApplyFit(eng).run(psi, 2, 20)
    f2 = Q.ApplyFit(eng)
    f2.set_option("graph", 0)
    b = f2.run(psi, 2, 20)
    assert np.array_equal(a["fidelities"], b["fidelities"])
    assert np.array_equal(a["psi"], b["psi"])
