"""A wavefunction file handed from one code base to the other: the oracle
writes a TensorTree file (TensorTree::write, TensorTree_Implementation.h:
79-92), the C ABI reads it, and the engine's sweep on the state read matches
the oracle's on the state written."""
import numpy as np
import pytest

import oracle_lib as O
import qutree_amd as Q
from qutree_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_sweep_on_a_state_read_from_a_file():
    text, H = W.cmf8_model()
    print("case formats cmf8 file round trip + sweep", flush=True)
    om = O.OracleModel(text, H)
    psi = om.fill_random(21, delta_lowest=False)
    data = om.tensor_tree_bytes(psi)
    back, used = Q.tensor_tree_from_bytes(text, data)
    assert used == len(data)
    ref = om.derivative(back)
    eng = Q.Engine(text, H)
    eng.set_psi(back)
    eng.derivative()
    got = eng.get(0)
    rel = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(f"  derivative max rel {rel:.2e}", flush=True)
    assert rel < 1e-10
    # and back out: the engine's state written by the C ABI is the oracle's file
    assert Q.tensor_tree_bytes(text, eng.get(2)) == data
